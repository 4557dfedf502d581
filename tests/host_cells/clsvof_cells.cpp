// Host build of the CLSVOF kernels' per-cell code (lsmp_clsvof.h of the
// library): the loops of k_face_velocity and k_wlic_pass, one cell at a time,
// so that tests can compare the kernels' arithmetic and indexing with the
// oracle on the CPU.  Test infrastructure only; the library never links it.
#include <cstring>
#include <utility>
#include <vector>

#include "../../mass-preserving-level-set-method_amd/csrc/lsmp_clsvof.h"

using namespace lsmp;

// one cell of a WLIC pass: the kernels' update, or the one that loads its
// stencil up front
template <int D, bool CLAMP>
double cell(int stencil, const WlicLine<double> &L, const double *old, int p, double dt, double dd)
{
    return stencil ? wlic_cell_stencil<D, CLAMP>(L, old, p, dt, dd) : wlic_cell<D, CLAMP>(L, old, p, dt, dd);
}

extern "C" {

// face velocities of a prescribed field into interior arrays (x fastest)
void hc_face_velocity(int nx, int ny, int nz, int field, double xstart, double xend, double ystart, double yend,
                      double zstart, double zend, double dx, double dy, double dz, double ct, double omega,
                      double *uf, double *vf, double *wf)
{
    FaceVelArgs va;
    va.nx = nx; va.ny = ny; va.nz = nz; va.field = field;
    va.xstart = xstart; va.ystart = ystart; va.zstart = zstart;
    va.dx = dx; va.dy = dy; va.dz = dz; va.ct = ct; va.omega = omega;
    std::vector<double> tab(3 * (size_t)(nx + ny + nz));
    for (int t = 0; t < nx; ++t) face_table_entry(xstart, xend, dx, nx, t, &tab[3 * (size_t)t]);
    for (int t = 0; t < ny; ++t) face_table_entry(ystart, yend, dy, ny, t, &tab[3 * (size_t)(nx + t)]);
    if (field == 2)
        for (int t = 0; t < nz; ++t) face_table_entry(zstart, zend, dz, nz, t, &tab[3 * (size_t)(nx + ny + t)]);
    const size_t n = (size_t)nx * ny * nz;
    for (size_t q = 0; q < n; ++q) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
        face_velocity_cell(va, tab.data(), i, j, k, &uf[q], &vf[q], &wf[q]);
    }
}

// the first npass of vof_wlic's passes over an interior vof array, in place
// (n2 null in 2D); stencil selects the cell update that loads its stencil up
// front
void hc_vof_wlic_n(int dim, int nx, int ny, int nz, double dt, double dx, double dy, double dz, double *vof,
                   const double *old, const double *uf, const double *vf, const double *wf, const double *n0,
                   const double *n1, const double *n2, int stencil, int npass)
{
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> in(vof, vof + n), out(n);
    const double *vel[3] = {uf, vf, wf};
    const double dd[3] = {dx, dy, dz};
    for (int d = 0; d < dim && d < npass; ++d) {
        const bool clamp = d == dim - 1;
        for (size_t q = 0; q < n; ++q) {
            int p;
            if (d == 0) {
                const WlicLine<double> L = wlic_line<0>(in.data(), vel[0], n0, n1, n2, q, nx, ny, nz, &p);
                out[q] = clamp ? cell<0, true>(stencil, L, old, p, dt, dd[0]) : cell<0, false>(stencil, L, old, p, dt, dd[0]);
            } else if (d == 1) {
                const WlicLine<double> L = wlic_line<1>(in.data(), vel[1], n0, n1, n2, q, nx, ny, nz, &p);
                out[q] = clamp ? cell<1, true>(stencil, L, old, p, dt, dd[1]) : cell<1, false>(stencil, L, old, p, dt, dd[1]);
            } else {
                const WlicLine<double> L = wlic_line<2>(in.data(), vel[2], n0, n1, n2, q, nx, ny, nz, &p);
                out[q] = cell<2, true>(stencil, L, old, p, dt, dd[2]);
            }
        }
        std::swap(in, out);
    }
    std::memcpy(vof, in.data(), n * sizeof(double));
}

// vof_wlic's passes over an interior vof array, in place (n2 null in 2D)
void hc_vof_wlic(int dim, int nx, int ny, int nz, double dt, double dx, double dy, double dz, double *vof,
                 const double *old, const double *uf, const double *vf, const double *wf, const double *n0,
                 const double *n1, const double *n2, int stencil)
{
    hc_vof_wlic_n(dim, nx, ny, nz, dt, dx, dy, dz, vof, old, uf, vf, wf, n0, n1, n2, stencil, dim);
}

// k_cut_line_flags on the host: flags of the x lines (ny nz), y lines (nx nz)
// and z lines (nx ny) through the cells wlic_may_cut selects
void hc_cut_line_flags(const double *vof, int nx, int ny, int nz, unsigned char *fx, unsigned char *fy,
                       unsigned char *fz)
{
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (!wlic_may_cut(vof, i, j, k, nx, ny, nz)) continue;
                fx[j + (size_t)ny * k] = 1;
                fy[i + (size_t)nx * k] = 1;
                if (nz > 1) fz[i + (size_t)nx * j] = 1;
            }
}

}  // extern "C"
