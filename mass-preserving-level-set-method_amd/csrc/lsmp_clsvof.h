// lsmp_clsvof.h -- per-cell arithmetic of the CLSVOF path: the face
// velocities of the prescribed fields and the WLIC-THINC VOF fluxes and
// update.  The gfx950 kernels (lsmp_clsvof.hip) call these functions, and a
// host build of the same header checks them cell by cell.
//
// Layout: interior-only arrays, x fastest (a = i + nx (j + ny k), 0-based).
// The reference's zero-gradient ghosts are index clamps.  Face p of a line
// (p = -1 .. n - 1) is the face between cells p and p + 1 (the reference's
// face i + 1/2 at index i, Fortran i = p + 1); its velocity is stored at cell
// clamp(p), so face -1 takes face 0's value as the reference's ghost fill of
// vel%x%old does.
#pragma once
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define LSMP_HD __host__ __device__ inline
#else
#include <cmath>
#include <cstddef>
#define LSMP_HD inline
#endif

namespace lsmp {

LSMP_HD int clamp_index(int p, int n) { return p < 0 ? 0 : p > n - 1 ? n - 1 : p; }

// One line of a WLIC pass: the cells base + p s, p = 0 .. n - 1, along the
// pass's direction.  n2 is null in 2D (the z normal is zero there).
template <typename T> struct WlicLine {
    const T *vof, *vel, *n0, *n1, *n2;
    size_t base, s;
    int n;
    LSMP_HD size_t at(int p) const { return base + (size_t)clamp_index(p, n) * s; }
};

// The line of direction D through interior cell q of an nx x ny x nz grid,
// and q's position *p on it.
template <int D, typename T>
LSMP_HD WlicLine<T> wlic_line(const T *vof, const T *vel, const T *n0, const T *n1, const T *n2, size_t q, int nx,
                              int ny, int nz, int *p)
{
    const size_t nxy = (size_t)nx * ny;
    int i, j, k;
    if (nxy * (size_t)nz <= 0xffffffffull) {
        // 32-bit quotients: a 64-bit division is a long integer sequence on the GPU
        const unsigned u = (unsigned)q, r = u / (unsigned)nx;
        i = (int)(u - r * (unsigned)nx);
        k = (int)(r / (unsigned)ny);
        j = (int)(r - (unsigned)k * (unsigned)ny);
    } else {
        i = (int)(q % nx);
        j = (int)((q / nx) % ny);
        k = (int)(q / nxy);
    }
    WlicLine<T> L;
    L.vof = vof;
    L.vel = vel;
    L.n0 = n0;
    L.n1 = n1;
    L.n2 = n2;
    if (D == 0) {
        L.base = q - (size_t)i;
        L.s = 1;
        L.n = nx;
        *p = i;
    } else if (D == 1) {
        L.base = (size_t)i + nxy * k;
        L.s = (size_t)nx;
        L.n = ny;
        *p = j;
    } else {
        L.base = (size_t)i + (size_t)nx * j;
        L.s = nxy;
        L.n = nz;
        *p = k;
    }
    return L;
}

// WLIC-THINC flux through face p of the line (clsvof.f90:26-63 for x,
// :91-128 for y, :191-228 for z).  The cell upwind of the face velocity gives
// vof; a full or empty cell is carried by the face velocity, a cut cell by
// the THINC profile 0.5 (1 + alpha tanh(beta (xi - xc))) (beta = 2.3, slope
// alpha from the cells either side of it, ib = max(1, ii - 1)), weighted with
// the upwind flux by |n_d| / (|n_x| + |n_y| + |n_z|).  D is the line's
// direction.  The arithmetic is the reference's, in its order.
template <int D, typename T> LSMP_HD double wlic_face_flux(const WlicLine<T> &L, int p, double dt, double dd)
{
    const double eps = 1.0e-10, beta = 2.3;
    const double uf = (double)L.vel[L.at(p)];
    int c;
    double isgn;
    if (uf > 0.0) {
        c = p;
        isgn = 1.0;
    } else {
        c = p + 1;
        isgn = 0.0;
    }
    const size_t qc = L.at(c);
    const double v = (double)L.vof[qc];
    if (v > 1.0 - eps || v < eps) return v * uf * dt;
    const int b = c - 1 > 0 ? c - 1 : 0;
    const double alpha = (double)L.vof[L.at(b)] < (double)L.vof[L.at(c + 1)] ? 1.0 : -1.0;
    const double a1 = exp(beta * (2.0 * v - 1.0) / alpha);
    const double a3 = exp(beta);
    const double xc = 0.5 / beta * log((a3 * a3 - a1 * a3) / (a1 * a3 - 1.0));
    const double a4 = cosh(beta * (isgn - uf * dt / dd - xc));
    const double a5 = cosh(beta * (isgn - xc));
    const double m0 = fabs((double)L.n0[qc]), m1 = fabs((double)L.n1[qc]);
    const double m2 = L.n2 ? fabs((double)L.n2[qc]) : 0.0;
    double w = m0 + m1 + m2;
    w = (D == 0 ? m0 : D == 1 ? m1 : m2) / w;
    double f = 0.5 * (uf * dt - alpha * dd / beta * log(a4 / a5));
    f = f * w + (1.0 - w) * v * uf * dt;
    return f;
}

// The update of cell p of the line by one direction's pass
// (clsvof.f90:74-79, :157-162, :251-256): the flux difference and the
// divergence term vof_old dt du/dd; CLAMP (the last direction) clips the
// result to [0, 1].
template <int D, bool CLAMP, typename T>
LSMP_HD double wlic_cell(const WlicLine<T> &L, const T *old, int p, double dt, double dd)
{
    const double fr = wlic_face_flux<D>(L, p, dt, dd);
    const double fl = wlic_face_flux<D>(L, p - 1, dt, dd);
    const size_t q = L.at(p);
    const double sflux = (fr - fl) / dd;
    double v = (double)L.vof[q] - sflux + (double)old[q] * dt * ((double)L.vel[q] - (double)L.vel[L.at(p - 1)]) / dd;
    if (CLAMP) v = fmin(1.0, fmax(0.0, v));
    return v;
}

// The same flux from values loaded beforehand: uf the face velocity and
// w[0..3] the vof of the cells p - 1 .. p + 2 around face p (clamped to the
// line).  The upwind cell's neighbour below, ib = max(1, ii - 1), is its
// clamped left neighbour: the zero-gradient ghosts equal the first cell.
// Only a cut cell's normals are loaded here.
template <int D, typename T>
LSMP_HD double wlic_face_flux_pre(const WlicLine<T> &L, int p, double uf, const double *w, double dt, double dd)
{
    const double eps = 1.0e-10, beta = 2.3;
    const bool up = uf > 0.0;
    const int c = up ? p : p + 1;
    const double isgn = up ? 1.0 : 0.0;
    const double v = up ? w[1] : w[2];
    if (v > 1.0 - eps || v < eps) return v * uf * dt;
    const double vb = up ? w[0] : w[1], va = up ? w[2] : w[3];
    const double alpha = vb < va ? 1.0 : -1.0;
    const double a1 = exp(beta * (2.0 * v - 1.0) / alpha);
    const double a3 = exp(beta);
    const double xc = 0.5 / beta * log((a3 * a3 - a1 * a3) / (a1 * a3 - 1.0));
    const double a4 = cosh(beta * (isgn - uf * dt / dd - xc));
    const double a5 = cosh(beta * (isgn - xc));
    const size_t qc = L.at(c);
    const double m0 = fabs((double)L.n0[qc]), m1 = fabs((double)L.n1[qc]);
    const double m2 = L.n2 ? fabs((double)L.n2[qc]) : 0.0;
    double wt = m0 + m1 + m2;
    wt = (D == 0 ? m0 : D == 1 ? m1 : m2) / wt;
    double f = 0.5 * (uf * dt - alpha * dd / beta * log(a4 / a5));
    f = f * wt + (1.0 - wt) * v * uf * dt;
    return f;
}

// wlic_cell with every load that does not depend on a cut cell issued up
// front (vof at p - 2 .. p + 2, the two face velocities, vof_old): one memory
// round trip per cell instead of a velocity load followed by the upwind vof
// load for each face.  Same operations on the same values as wlic_cell.
template <int D, bool CLAMP, typename T>
LSMP_HD double wlic_cell_stencil(const WlicLine<T> &L, const T *old, int p, double dt, double dd)
{
    double w[5];
    for (int o = 0; o < 5; ++o) w[o] = (double)L.vof[L.at(p - 2 + o)];
    const double ul = (double)L.vel[L.at(p - 1)], ur = (double)L.vel[L.at(p)];
    const double vo = (double)old[L.at(p)];
    const double fr = wlic_face_flux_pre<D>(L, p, ur, w + 1, dt, dd);
    const double fl = wlic_face_flux_pre<D>(L, p - 1, ul, w, dt, dd);
    const double sflux = (fr - fl) / dd;
    double v = w[2] - sflux + vo * dt * (ur - ul) / dd;
    if (CLAMP) v = fmin(1.0, fmax(0.0, v));
    return v;
}

// Whether a WLIC step can find cell (i, j, k) cut in any of its passes: not
// when every cell of its 3 x 3 x 3 neighbourhood (clamped) is empty (vof <=
// 1e-11) or every one is full (vof >= 1 - 1e-11).  Each pass moves vof by at
// most one cell along its direction (|u| dt < dd), so after the x and y passes
// a cell has received vof only from cells up to one step away per axis; and
// near-empty or near-full neighbourhoods change by less than their 1e-11
// distance from 0 or 1 times a few, short of the flux's 1e-10 cut threshold.
template <typename T> LSMP_HD bool wlic_may_cut(const T *vof, int i, int j, int k, int nx, int ny, int nz)
{
    double lo = 2.0, hi = -1.0;
    for (int dk = -1; dk <= 1; ++dk) {
        const size_t zk = (size_t)clamp_index(k + dk, nz) * (size_t)ny;
        for (int dj = -1; dj <= 1; ++dj) {
            const size_t row = ((size_t)clamp_index(j + dj, ny) + zk) * (size_t)nx;
            for (int di = -1; di <= 1; ++di) {
                const double v = (double)vof[row + (size_t)clamp_index(i + di, nx)];
                lo = fmin(lo, v);
                hi = fmax(hi, v);
            }
        }
    }
    return hi > 1.0e-11 && lo < 1.0 - 1.0e-11;
}

// Face velocities of the prescribed fields.  Per axis and cell the table holds
// sin(pi c), sin(2 pi c) at the cell centre c and sin(pi f) at the face
// midpoint f = (c + c') / 2, c' the next cell's centre (the wall for the last
// cell, as the reference's x(ie + 1) = xend).
struct FaceVelArgs {
    int nx, ny, nz, field;
    double xstart, ystart, zstart, dx, dy, dz, ct, omega;
};

LSMP_HD double sqd(double a) { return a * a; }

// (Disk/src/setup/velocity.f90:89-93: the rotation's faces take the
// cell-centre formulas; Vortex2D/src/setup/velocity.f90:38-39;
// Vortex/src/setup/velocity.f90:31-33)
LSMP_HD void face_velocity_cell(const FaceVelArgs &va, const double *tab, int i, int j, int k, double *uu,
                                double *vv, double *ww)
{
    if (va.field == 0) {
        const double x = va.xstart + (double)((float)(i + 1) - 0.5f) * va.dx;
        const double y = va.ystart + (double)((float)(j + 1) - 0.5f) * va.dy;
        *uu = -va.omega * (y - 0.5);
        *vv = va.omega * (x - 0.5);
        *ww = 0.0;
        return;
    }
    const double *tx = tab, *ty = tab + 3 * va.nx, *tz = ty + 3 * va.ny;
    const double s2x = tx[3 * i + 1], sfx = tx[3 * i + 2];
    const double s2y = ty[3 * j + 1], sfy = ty[3 * j + 2];
    if (va.field == 1) {
        *uu = -sqd(sfx) * s2y * va.ct;
        *vv = sqd(sfy) * s2x * va.ct;
        *ww = 0.0;
        return;
    }
    const double s2z = tz[3 * k + 1], sfz = tz[3 * k + 2];
    *uu = 2.0 * sqd(sfx) * s2y * s2z * va.ct;
    *vv = -s2x * sqd(sfy) * s2z * va.ct;
    *ww = -s2x * s2y * sqd(sfz) * va.ct;
}

// Table entry t of face_velocity_cell's table: axis cell index a of n cells
// starting at start with spacing d, end the axis's far wall.
LSMP_HD void face_table_entry(double start, double end, double d, int n, int a, double *out3)
{
    const double pi = 3.141592653589793238462643383279502884;
    const double c = start + (double)((float)(a + 1) - 0.5f) * d;
    const double cn = a + 1 < n ? start + (double)((float)(a + 2) - 0.5f) * d : end;
    const double f = 0.5 * (c + cn);
    out3[0] = sin(pi * c);
    out3[1] = sin(2.0 * pi * c);
    out3[2] = sin(pi * f);
}

}  // namespace lsmp
