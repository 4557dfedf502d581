// lsmp_clsvof.hip -- gfx950 kernels of the CLSVOF path (coupled level set and
// volume of fluid, Vortex/src/interface_solver/clsvof.f90): the prescribed
// face velocities, the three WLIC-THINC passes and the elementwise steps of
// the reconstruction.  The CCD normals and the reinitialisation are the
// level-set path's kernels.
//
// Every kernel is elementwise over the interior cells, one thread per cell:
// a WLIC pass reads vof along its direction (cells p - 2 .. p + 2, served by
// L1/L2), vof_old, the face velocity, and the three normals only on the cut
// cells; it writes one value.  Workgroups are numbered XCD-contiguously, so a
// y or z pass finds the neighbouring rows in the same L2.
#include <hip/hip_runtime.h>

#include "lsmp_clsvof.h"
#include "lsmp_device.h"
#include "lsmp_internal.h"

namespace lsmp {

namespace {

inline unsigned nblocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// per-axis table of face_velocity_cell: nx + ny (+ nz) entries of 3 values
__global__ void k_face_tables(FaceVelArgs va, double xend, double yend, double zend, double *__restrict__ tab)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = va.nx + va.ny + (va.field == 2 ? va.nz : 0);
    if (t >= nt) return;
    if (t < va.nx) face_table_entry(va.xstart, xend, va.dx, va.nx, t, tab + 3 * t);
    else if (t < va.nx + va.ny) face_table_entry(va.ystart, yend, va.dy, va.ny, t - va.nx, tab + 3 * t);
    else face_table_entry(va.zstart, zend, va.dz, va.nz, t - va.nx - va.ny, tab + 3 * t);
}

template <typename T>
__global__ void k_face_velocity(T *__restrict__ uf, T *__restrict__ vf, T *__restrict__ wf, FaceVelArgs va,
                                const double *__restrict__ tab)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)va.nx * va.ny * va.nz;
    if (q >= n) return;
    const int i = (int)(q % va.nx), j = (int)((q / va.nx) % va.ny), k = (int)(q / ((size_t)va.nx * va.ny));
    double uu, vv, ww;
    face_velocity_cell(va, tab, i, j, k, &uu, &vv, &ww);
    uf[q] = (T)uu;
    vf[q] = (T)vv;
    wf[q] = (T)ww;
}

// One direction's WLIC pass: out = the updated vof of every cell from in
// (vof_wlic_solver's flux loop and update loop of direction D), with the cell
// update that issues its stencil's loads up front (wlic_cell_stencil: the
// operations of wlic_cell on the same values; 17% faster at 256^3).
template <typename T, int D, bool CLAMP>
__global__ void __launch_bounds__(256) k_wlic_pass(const T *__restrict__ in, T *__restrict__ out,
                                                   const T *__restrict__ old, const T *__restrict__ vel,
                                                   const T *__restrict__ n0, const T *__restrict__ n1,
                                                   const T *__restrict__ n2, int nx, int ny, int nz, double dt,
                                                   double dd)
{
    const size_t q = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    if (q >= (size_t)nx * ny * nz) return;
    int p;
    const WlicLine<T> L = wlic_line<D>(in, vel, n0, n1, n2, q, nx, ny, nz, &p);
    out[q] = (T)wlic_cell_stencil<D, CLAMP>(L, old, p, dt, dd);
}

// Flags of the lines through each cell a WLIC step can find cut (wlic_may_cut).
template <typename T>
__global__ void k_cut_line_flags(const T *__restrict__ vof, int nx, int ny, int nz, unsigned char *__restrict__ fx,
                                 unsigned char *__restrict__ fy, unsigned char *__restrict__ fz)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)nx * ny * nz) return;
    const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
    if (!wlic_may_cut(vof, i, j, k, nx, ny, nz)) return;
    fx[j + (size_t)ny * k] = 1;
    fy[i + (size_t)nx * k] = 1;
    if (nz > 1) fz[i + (size_t)nx * j] = 1;
}

// clsvof_recon, first loop (clsvof.f90:283-292): keep phi, phi <- 2 vof - 1
template <typename T>
__global__ void k_recon_start(T *__restrict__ phi, const T *__restrict__ vof, T *__restrict__ keep, size_t n)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    keep[q] = phi[q];
    phi[q] = T(2) * vof[q] - T(1);
}

// clsvof_recon, second loop (:305-311): the kept phi where |phi_kept| > w
template <typename T>
__global__ void k_recon_keep(T *__restrict__ phi, const T *__restrict__ keep, size_t n, double w)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const T k = keep[q];
    if (fabs((double)k) > w) phi[q] = k;
}

// Block partial sums over the VOF: l1 == 0, the VOF half of ls_mv
// (ns_solver/others.f90:215-247): sum vof dv and sum rho(vof) vof dv; l1 != 0,
// the VOF part of calculate_interface_loss (output.f90:178-204): sum |vof0 - vof|.
template <typename T>
__global__ void k_vof_sums(const T *__restrict__ vof, const T *__restrict__ vof0, size_t n, double dv, double r12,
                           int inverse, int l1, double *__restrict__ partial)
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double v = (double)vof[q];
        if (l1) {
            acc[0] += fabs((double)vof0[q] - v);
        } else {
            acc[0] += v * dv;
            const double rho = inverse ? r12 * v + (1.0 - v) : v + r12 * (1.0 - v);
            acc[1] += rho * v * dv;
        }
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = acc[0];
        partial[3 * blockIdx.x + 1] = acc[1];
        partial[3 * blockIdx.x + 2] = 0.0;
    }
}

}  // namespace

template <typename T>
void launch_vof_sums(hipStream_t st, const T *vof, const T *vof0, size_t n, double dv, double r12, int inverse,
                     int l1, double *partial, int nblocks)
{
    hipLaunchKernelGGL((k_vof_sums<T>), dim3(nblocks), dim3(256), 0, st, vof, vof0, n, dv, r12, inverse, l1,
                       partial);
}

template <typename T>
void launch_face_velocity(hipStream_t st, T *uf, T *vf, T *wf, const FaceVelArgs &va, double xend, double yend,
                          double zend, double *tab)
{
    const size_t n = (size_t)va.nx * va.ny * va.nz;
    if (va.field != 0) {
        const int nt = va.nx + va.ny + (va.field == 2 ? va.nz : 0);
        hipLaunchKernelGGL(k_face_tables, dim3(nblocks(nt, 256)), dim3(256), 0, st, va, xend, yend, zend, tab);
    }
    hipLaunchKernelGGL((k_face_velocity<T>), dim3(nblocks(n, 256)), dim3(256), 0, st, uf, vf, wf, va,
                       (const double *)tab);
}

template <typename T>
void launch_wlic_pass(hipStream_t st, int dir, bool clamp, const T *in, T *out, const T *old, const T *vel,
                      const T *n0, const T *n1, const T *n2, int nx, int ny, int nz, double dt, double dd)
{
    const size_t n = (size_t)nx * ny * nz;
    const dim3 g(nblocks(n, 256)), b(256);
#define LSMP_WLIC(D, C)                                                                                        \
    hipLaunchKernelGGL((k_wlic_pass<T, D, C>), g, b, 0, st, in, out, old, vel, n0, n1, n2, nx, ny, nz, dt, dd)
    if (dir == 0) {
        if (clamp) LSMP_WLIC(0, true);
        else LSMP_WLIC(0, false);
    } else if (dir == 1) {
        if (clamp) LSMP_WLIC(1, true);
        else LSMP_WLIC(1, false);
    } else {
        LSMP_WLIC(2, true);
    }
#undef LSMP_WLIC
}

template <typename T>
void launch_cut_lines(hipStream_t st, int dim, const T *vof, int nx, int ny, int nz, unsigned char *flags,
                      unsigned *count, int *listx, int *listy, int *listz)
{
    const int nfx = ny * nz, nfy = nx * nz, nfz = dim == 3 ? nx * ny : 0;
    launch_zero_bytes(st, flags, band_flag_bytes(nfx, nfy, nfz));
    unsigned char *fy = flags + band_flag_offset(1, nfx, nfy), *fz = flags + band_flag_offset(2, nfx, nfy);
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL((k_cut_line_flags<T>), dim3(nblocks(n, 256)), dim3(256), 0, st, vof, nx, ny, nz, flags, fy,
                       fz);
    launch_compact_lines(st, dim, flags, nfx, nfy, nfz, listx, listy, listz, count);
}

template <typename T> void launch_recon_start(hipStream_t st, T *phi, const T *vof, T *keep, size_t n)
{
    hipLaunchKernelGGL((k_recon_start<T>), dim3(nblocks(n, 256)), dim3(256), 0, st, phi, vof, keep, n);
}

template <typename T> void launch_recon_keep(hipStream_t st, T *phi, const T *keep, size_t n, double w)
{
    hipLaunchKernelGGL((k_recon_keep<T>), dim3(nblocks(n, 256)), dim3(256), 0, st, phi, keep, n, w);
}

#define LSMP_CLSVOF_INST(T)                                                                                    \
    template void launch_face_velocity<T>(hipStream_t, T *, T *, T *, const FaceVelArgs &, double, double,     \
                                          double, double *);                                                   \
    template void launch_wlic_pass<T>(hipStream_t, int, bool, const T *, T *, const T *, const T *, const T *, \
                                      const T *, const T *, int, int, int, double, double);                    \
    template void launch_recon_start<T>(hipStream_t, T *, const T *, T *, size_t);                             \
    template void launch_cut_lines<T>(hipStream_t, int, const T *, int, int, int, unsigned char *, unsigned *,  \
                                      int *, int *, int *);                                                    \
    template void launch_recon_keep<T>(hipStream_t, T *, const T *, size_t, double);                          \
    template void launch_vof_sums<T>(hipStream_t, const T *, const T *, size_t, double, double, int, int,      \
                                     double *, int);

LSMP_CLSVOF_INST(double)
LSMP_CLSVOF_INST(float)

}  // namespace lsmp
