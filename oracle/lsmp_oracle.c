/*
 * lsmp_oracle.c -- scalar FP64 CPU restatement of the reference's MPLS
 * interface-transport step.  TEST INFRASTRUCTURE ONLY (see lsmp_oracle.h).
 *
 * Compile with -O2 -ffp-contract=off so that every expression rounds as the
 * reference's Fortran does when built without FMA contraction.  Expressions
 * keep the reference's association order; Fortran evaluates a*b/c*d left to
 * right, and x**2.0d0 is a square (both spelled out below).
 *
 * Loops follow the reference's k, j, i (i fastest) order so that the
 * sequential sums reproduce an OpenMP build run on one thread.
 *
 * Built with -fopenmp (the CPU baseline's build, never the checker's), the
 * loops whose iterations are independent -- pointwise updates, line solves,
 * ghost fills -- and the max reduction of the pseudo-time stop run on every
 * host thread; the sums stay sequential, so both builds give the same
 * results.
 */
#include "lsmp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GHC 3

#ifdef _OPENMP
#include <omp.h>
#define ORC_OMP(x) _Pragma(x)
static int orc_threads(void) { return omp_get_max_threads(); }
static int orc_tid(void) { return omp_get_thread_num(); }
#else
#define ORC_OMP(x)
static int orc_threads(void) { return 1; }
static int orc_tid(void) { return 0; }
#endif
#define ORC_PAR ORC_OMP("omp parallel for collapse(2) schedule(static)")

struct orc_state {
    orc_params p;
    int nxb, nyb, nzb;      /* box extents */
    int ks, ke;             /* interior z range (1..nz) */
    long nbox;
    double *f[ORC_NFIELDS];
    /* scalars (glb) */
    double mass, vol, massv, volv, imass, ivol, red_error;
    double imass1, imass2;  /* marker(1:2)%imass */
    int merged;             /* glb%merged */
    double time;
    int iter;
    int last_redis_iters, last_mass_iters;
    int scheme;             /* advection RHS: 0 CRWENO flux split, 1 UCCD */
};

static inline double sq(double x) { return x * x; }

/* box index for Fortran indices (i, j, k), i in 1-GHC .. n+GHC */
static inline long BX(const orc_state *s, int i, int j, int k)
{
    return ((long)(k - 1 + GHC) * s->nyb + (j - 1 + GHC)) * s->nxb + (i - 1 + GHC);
}
#define AT(arr, i, j, k) (arr)[BX(s, (i), (j), (k))]

orc_state *orc_create(const orc_params *p)
{
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *p;
    s->nxb = p->nx + 2 * GHC;
    s->nyb = p->ny + 2 * GHC;
    if (p->dim == 2) {
        s->p.nz = 1;
        s->nzb = 1 + 2 * GHC;
    } else {
        s->nzb = p->nz + 2 * GHC;
    }
    s->ks = 1;
    s->ke = s->p.nz;
    s->nbox = (long)s->nxb * s->nyb * s->nzb;
    for (int q = 0; q < ORC_NFIELDS; ++q)
        s->f[q] = (double *)calloc((size_t)s->nbox, sizeof(double));
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    for (int q = 0; q < ORC_NFIELDS; ++q) free(s->f[q]);
    free(s);
}

double *orc_field(orc_state *s, int which) { return s->f[which]; }
long orc_box_size(const orc_state *s) { return s->nbox; }
void orc_box_dims(const orc_state *s, int *nxb, int *nyb, int *nzb)
{
    *nxb = s->nxb; *nyb = s->nyb; *nzb = s->nzb;
}

double orc_get(const orc_state *s, const char *n)
{
    if (!strcmp(n, "mass")) return s->mass;
    if (!strcmp(n, "vol")) return s->vol;
    if (!strcmp(n, "massv")) return s->massv;
    if (!strcmp(n, "volv")) return s->volv;
    if (!strcmp(n, "imass")) return s->imass;
    if (!strcmp(n, "imass1")) return s->imass1;
    if (!strcmp(n, "imass2")) return s->imass2;
    if (!strcmp(n, "merged")) return s->merged;
    if (!strcmp(n, "ivol")) return s->ivol;
    if (!strcmp(n, "red_error")) return s->red_error;
    if (!strcmp(n, "time")) return s->time;
    if (!strcmp(n, "iter")) return (double)s->iter;
    if (!strcmp(n, "dt")) return s->p.dt;
    if (!strcmp(n, "redis_iters")) return (double)s->last_redis_iters;
    if (!strcmp(n, "mass_iters")) return (double)s->last_mass_iters;
    if (!strcmp(n, "scheme")) return (double)s->scheme;
    return NAN;
}

void orc_set(orc_state *s, const char *n, double v)
{
    if (!strcmp(n, "imass")) s->imass = v;
    if (!strcmp(n, "imass1")) s->imass1 = v;
    if (!strcmp(n, "imass2")) s->imass2 = v;
    if (!strcmp(n, "merged")) s->merged = (int)v;
    else if (!strcmp(n, "ivol")) s->ivol = v;
    else if (!strcmp(n, "time")) s->time = v;
    else if (!strcmp(n, "iter")) s->iter = (int)v;
    else if (!strcmp(n, "dt")) s->p.dt = v;
    else if (!strcmp(n, "mass")) s->mass = v;
    else if (!strcmp(n, "scheme")) s->scheme = (int)v;
}

/* ------------------------------------------------------------------------
 * job_bc order 0: zero-gradient ghost copy, x faces over the whole (j,k) box,
 * then y over the whole (i,k) box, then z (Disk/lib/data/branches.f90:283-310).
 * ---------------------------------------------------------------------- */
void orc_bc(orc_state *s, double *dat)
{
    const int nx = s->p.nx, ny = s->p.ny;
    const int ks = s->ks, ke = s->ke;
    ORC_PAR
    for (int k = ks - GHC; k <= ke + GHC; ++k)
        for (int j = 1 - GHC; j <= ny + GHC; ++j)
            for (int i = 1; i <= GHC; ++i) {
                AT(dat, 1 - i, j, k) = AT(dat, 1, j, k);
                AT(dat, nx + i, j, k) = AT(dat, nx, j, k);
            }
    ORC_PAR
    for (int k = ks - GHC; k <= ke + GHC; ++k)
        for (int i = 1 - GHC; i <= nx + GHC; ++i)
            for (int j = 1; j <= GHC; ++j) {
                AT(dat, i, 1 - j, k) = AT(dat, i, 1, k);
                AT(dat, i, ny + j, k) = AT(dat, i, ny, k);
            }
    ORC_PAR
    for (int j = 1 - GHC; j <= ny + GHC; ++j)
        for (int i = 1 - GHC; i <= nx + GHC; ++i)
            for (int k = 1; k <= GHC; ++k) {
                AT(dat, i, j, ks - k) = AT(dat, i, j, ks);
                AT(dat, i, j, ke + k) = AT(dat, i, j, ke);
            }
}

/* ------------------------------------------------------------------------
 * ls_funs: smoothed Heaviside, delta and sign on the full box
 * (Vortex/src/ns_solver/others.f90:1-45).
 * ---------------------------------------------------------------------- */
void orc_ls_funs(orc_state *s)
{
    const double eps = 1.0e-12;
    const double pi = acos(-1.0);
    const double w = s->p.ls_wid;
    double *phi = s->f[ORC_PHI], *H = s->f[ORC_HEAVY], *D = s->f[ORC_DELTA], *SG = s->f[ORC_SIGN];
    ORC_OMP("omp parallel for schedule(static)")
    for (long q = 0; q < s->nbox; ++q) {
        double x = phi[q] / w;
        double heavy, hp;
        if (x > 1.0 - eps) {
            heavy = 1.0; hp = 0.0;
        } else if (x < -1.0 + eps) {
            heavy = 0.0; hp = 0.0;
        } else {
            heavy = 0.5 * ((1.0 + x) + sin(pi * x) / pi);
            hp = 0.5 * (1.0 + cos(pi * x)) / w;
        }
        heavy = fmax(fmin(heavy, 1.0), 0.0);
        H[q] = heavy;
        D[q] = hp;
        SG[q] = 2.0 * heavy - 1.0;
    }
}

/* ------------------------------------------------------------------------
 * ls_mv: LS and VOF mass / volume sums over the interior
 * (Vortex/src/ns_solver/others.f90:172-243).
 * ---------------------------------------------------------------------- */
void orc_ls_mv(orc_state *s)
{
    const double dv = s->p.dx * s->p.dy * s->p.dz;
    const double r12 = s->p.rho_12;
    orc_ls_funs(s);
    double *H = s->f[ORC_HEAVY], *V = s->f[ORC_VOF];
    double mass = 0.0, vol = 0.0;
    for (int k = s->ks; k <= s->ke; ++k)
        for (int j = 1; j <= s->p.ny; ++j)
            for (int i = 1; i <= s->p.nx; ++i) {
                double h = AT(H, i, j, k), rho;
                vol = vol + h * dv;
                if (!s->p.inverse) rho = h + r12 * (1.0 - h);
                else rho = r12 * h + (1.0 - h);
                mass = mass + rho * h * dv;
            }
    s->mass = mass;
    s->vol = vol;
    mass = 0.0; vol = 0.0;
    for (int k = s->ks; k <= s->ke; ++k)
        for (int j = 1; j <= s->p.ny; ++j)
            for (int i = 1; i <= s->p.nx; ++i) {
                double h = AT(V, i, j, k), rho;
                vol = vol + h * dv;
                if (!s->p.inverse) rho = h + r12 * (1.0 - h);
                else rho = r12 * h + (1.0 - h);
                mass = mass + rho * h * dv;
            }
    s->massv = mass;
    s->volv = vol;
}

/* ------------------------------------------------------------------------
 * Thomas algorithm without pivoting (Vortex/src/others/TDMA.f90:1-36).
 * ---------------------------------------------------------------------- */
void orc_solve_tridiagonal(double *A, double *B, double *C, double *S, double *X, int n)
{
    C[0] = C[0] / B[0];
    S[0] = S[0] / B[0];
    for (int i = 1; i < n; ++i) {
        C[i] = C[i] / (B[i] - A[i] * C[i - 1]);
        S[i] = (S[i] - A[i] * S[i - 1]) / (B[i] - A[i] * C[i - 1]);
    }
    X[n - 1] = S[n - 1];
    for (int i = n - 2; i >= 0; --i) X[i] = S[i] - C[i] * X[i + 1];
}

/* line arrays: Fortran index i (is-ghc .. ie+ghc, is = 1) stored at [i + 2] */
#define L(a, i) (a)[(i) + 2]

/* WENO-JS5 split reconstruction on the whole line; only the two boundary
 * interfaces survive CRWENO (level_set_rk3_solver.f90:288-345). */
void orc_wenojs_flux_split(const double *f, const double *g, double *fp, double *gm, int n)
{
    const double EPS = 1.0e-13;
    const int is = 1, ie = n;
    for (int i = is - GHC + 2; i <= ie + GHC - 3; ++i) {
        double b1 = 13.0 * sq(L(g, i - 2) - 2.0 * L(g, i - 1) + L(g, i)) +
                    3.0 * sq(L(g, i - 2) - 4.0 * L(g, i - 1) + 3.0 * L(g, i));
        double b2 = 13.0 * sq(L(g, i - 1) - 2.0 * L(g, i) + L(g, i + 1)) +
                    3.0 * sq(L(g, i - 1) - L(g, i + 1));
        double b3 = 13.0 * sq(L(g, i) - 2.0 * L(g, i + 1) + L(g, i + 2)) +
                    3.0 * sq(3.0 * L(g, i) - 4.0 * L(g, i + 1) + L(g, i + 2));
        double a1 = 1.0 / sq(EPS + b1);
        double a2 = 6.0 / sq(EPS + b2);
        double a3 = 3.0 / sq(EPS + b3);
        double w1 = a1 / (a1 + a2 + a3);
        double w2 = a2 / (a1 + a2 + a3);
        double w3 = a3 / (a1 + a2 + a3);
        L(gm, i) = w1 / 3.0 * L(g, i - 2) - (7.0 * w1 + w2) / 6.0 * L(g, i - 1) +
                   (11.0 * w1 + 5.0 * w2 + 2.0 * w3) / 6.0 * L(g, i) +
                   (2.0 * w2 + 5.0 * w3) / 6.0 * L(g, i + 1) - w3 / 6.0 * L(g, i + 2);
    }
    for (int i = is - GHC + 2; i <= ie + GHC - 3; ++i) {
        double b3 = 13.0 * sq(L(f, i - 1) - 2.0 * L(f, i) + L(f, i + 1)) +
                    3.0 * sq(L(f, i - 1) - 4.0 * L(f, i) + 3.0 * L(f, i + 1));
        double b2 = 13.0 * sq(L(f, i) - 2.0 * L(f, i + 1) + L(f, i + 2)) +
                    3.0 * sq(L(f, i) - L(f, i + 2));
        double b1 = 13.0 * sq(L(f, i + 1) - 2.0 * L(f, i + 2) + L(f, i + 3)) +
                    3.0 * sq(3.0 * L(f, i + 1) - 4.0 * L(f, i + 2) + L(f, i + 3));
        double a1 = 1.0 / sq(EPS + b1);
        double a2 = 6.0 / sq(EPS + b2);
        double a3 = 3.0 / sq(EPS + b3);
        double w1 = a1 / (a1 + a2 + a3);
        double w2 = a2 / (a1 + a2 + a3);
        double w3 = a3 / (a1 + a2 + a3);
        L(fp, i) = w3 * (-L(f, i - 1) + 5.0 * L(f, i) + 2.0 * L(f, i + 1)) / 6.0 +
                   w2 * (2.0 * L(f, i) + 5.0 * L(f, i + 1) - L(f, i + 2)) / 6.0 +
                   w1 * (11.0 * L(f, i + 1) - 7.0 * L(f, i + 2) + 2.0 * L(f, i + 3)) / 6.0;
    }
}

/* CRWENO5 with WENO-Z-type weights; two tridiagonal systems per line
 * (level_set_rk3_solver.f90:347-414). */
void orc_crweno_flux_split(const double *f, const double *g, double *fp, double *gm, int n)
{
    const double EPS = 1.0e-13;
    const double c1 = 0.2089141306, c2 = 0.4999999998, c3 = 0.2910858692;
    const int is = 1, ie = n;
    const int m0 = is - GHC + 3, m1 = ie + GHC - 4;   /* = 1 .. n-1 */
    const int nn = m1 - m0 + 1;
    double *A = (double *)malloc(sizeof(double) * (size_t)(nn > 0 ? nn : 1));
    double *B = (double *)malloc(sizeof(double) * (size_t)(nn > 0 ? nn : 1));
    double *C = (double *)malloc(sizeof(double) * (size_t)(nn > 0 ? nn : 1));
    double *S = (double *)malloc(sizeof(double) * (size_t)(nn > 0 ? nn : 1));

    orc_wenojs_flux_split(f, g, fp, gm, n);

    for (int i = m0; i <= m1; ++i) {
        double b1 = 13.0 * sq(L(g, i - 2) - 2.0 * L(g, i - 1) + L(g, i)) +
                    3.0 * sq(L(g, i - 2) - 4.0 * L(g, i - 1) + 3.0 * L(g, i));
        double b2 = 13.0 * sq(L(g, i - 1) - 2.0 * L(g, i) + L(g, i + 1)) +
                    3.0 * sq(L(g, i - 1) - L(g, i + 1));
        double b3 = 13.0 * sq(L(g, i) - 2.0 * L(g, i + 1) + L(g, i + 2)) +
                    3.0 * sq(3.0 * L(g, i) - 4.0 * L(g, i + 1) + L(g, i + 2));
        double a1 = c1 * (1.0 + fabs(b3 - b1) / (EPS + b1));
        double a2 = c2 * (1.0 + fabs(b3 - b1) / (EPS + b2));
        double a3 = c3 * (1.0 + fabs(b3 - b1) / (EPS + b3));
        double w1 = a1 / (a1 + a2 + a3);
        double w2 = a2 / (a1 + a2 + a3);
        double w3 = a3 / (a1 + a2 + a3);
        int r = i - m0;
        A[r] = (2.0 * w1 + w2) / 3.0;
        B[r] = (w1 + 2.0 * (w2 + w3)) / 3.0;
        C[r] = w3 / 3.0;
        S[r] = w1 / 6.0 * L(g, i - 1) + (5.0 * (w1 + w2) + w3) / 6.0 * L(g, i) +
               (w2 + 5.0 * w3) / 6.0 * L(g, i + 1);
    }
    if (nn > 0) {
        S[0] = S[0] - A[0] * L(gm, m0 - 1);
        S[nn - 1] = S[nn - 1] - C[nn - 1] * L(gm, m1 + 1);
        orc_solve_tridiagonal(A, B, C, S, &L(gm, m0), nn);
    }

    for (int i = m0; i <= m1; ++i) {
        double b3 = 13.0 * sq(L(f, i - 1) - 2.0 * L(f, i) + L(f, i + 1)) +
                    3.0 * sq(L(f, i - 1) - 4.0 * L(f, i) + 3.0 * L(f, i + 1));
        double b2 = 13.0 * sq(L(f, i) - 2.0 * L(f, i + 1) + L(f, i + 2)) +
                    3.0 * sq(L(f, i) - L(f, i + 2));
        double b1 = 13.0 * sq(L(f, i + 1) - 2.0 * L(f, i + 2) + L(f, i + 3)) +
                    3.0 * sq(3.0 * L(f, i + 1) - 4.0 * L(f, i + 2) + L(f, i + 3));
        double a1 = c1 * (1.0 + fabs(b3 - b1) / (EPS + b1));
        double a2 = c2 * (1.0 + fabs(b3 - b1) / (EPS + b2));
        double a3 = c3 * (1.0 + fabs(b3 - b1) / (EPS + b3));
        double w1 = a1 / (a1 + a2 + a3);
        double w2 = a2 / (a1 + a2 + a3);
        double w3 = a3 / (a1 + a2 + a3);
        int r = i - m0;
        A[r] = (w3) / 3.0;
        B[r] = (w1 + 2.0 * (w2 + w3)) / 3.0;
        C[r] = (w2 + 2.0 * w1) / 3.0;
        S[r] = (5.0 * w3 + w2) / 6.0 * L(f, i) + (w3 + 5.0 * (w2 + w1)) / 6.0 * L(f, i + 1) +
               w1 / 6.0 * L(f, i + 2);
    }
    if (nn > 0) {
        S[0] = S[0] - A[0] * L(fp, m0 - 1);
        S[nn - 1] = S[nn - 1] - C[nn - 1] * L(fp, m1 + 1);
        orc_solve_tridiagonal(A, B, C, S, &L(fp, m0), nn);
    }
    free(A); free(B); free(C); free(S);
}

/* ------------------------------------------------------------------------
 * Advection right-hand side (level_set_rk3_solver.f90:104-221; the 3D z
 * terms follow Vortex/, the 2D form Disk/ lines 166-174, 210-212).
 * ---------------------------------------------------------------------- */
void orc_rk3_source(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const int three_d = (s->p.dim == 3);
    double *phi = s->f[ORC_PHI];
    double *u = s->f[ORC_U], *v = s->f[ORC_V], *w = s->f[ORC_W];
    double *sx1 = s->f[ORC_SX1], *sx2 = s->f[ORC_SX2], *sy1 = s->f[ORC_SY1], *sy2 = s->f[ORC_SY2];
    double *sz1 = s->f[ORC_SZ1], *sz2 = s->f[ORC_SZ2];
    double *ssx1 = s->f[ORC_SSX1], *ssx2 = s->f[ORC_SSX2], *ssy1 = s->f[ORC_SSY1];
    double *ssy2 = s->f[ORC_SSY2], *ssz1 = s->f[ORC_SSZ1], *ssz2 = s->f[ORC_SSZ2];
    double *tmp = s->f[ORC_PHI_TMP];

    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double ph = AT(phi, i, j, k);
                double uu = AT(u, i, j, k), vv = AT(v, i, j, k), ww = AT(w, i, j, k);
                AT(sx1, i, j, k) = 0.5 * (uu + fabs(uu)) * ph;
                AT(sx2, i, j, k) = 0.5 * (uu - fabs(uu)) * ph;
                AT(sy1, i, j, k) = 0.5 * (vv + fabs(vv)) * ph;
                AT(sy2, i, j, k) = 0.5 * (vv - fabs(vv)) * ph;
                AT(sz1, i, j, k) = 0.5 * (ww + fabs(ww)) * ph;
                AT(sz2, i, j, k) = 0.5 * (ww - fabs(ww)) * ph;
            }
    orc_bc(s, sx1); orc_bc(s, sx2);
    orc_bc(s, sy1); orc_bc(s, sy2);
    orc_bc(s, sz1); orc_bc(s, sz2);

    const int nmax = (nx > ny ? nx : ny) > s->p.nz ? (nx > ny ? nx : ny) : s->p.nz;
    const int len = nmax + 2 * GHC;
    double *scr = (double *)malloc(sizeof(double) * 4 * len * orc_threads());
    ORC_OMP("omp parallel")
    {
    double *lf = scr + 4 * len * orc_tid(), *lg = lf + len, *lfp = lg + len, *lgm = lfp + len;

    /* x lines (:,j,k) */
    ORC_OMP("omp for collapse(2) schedule(static)")
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j) {
            for (int i = 1 - GHC; i <= nx + GHC; ++i) {
                L(lf, i) = AT(sx2, i, j, k); L(lg, i) = AT(sx1, i, j, k);
                L(lfp, i) = AT(ssx2, i, j, k); L(lgm, i) = AT(ssx1, i, j, k);
            }
            orc_crweno_flux_split(lf, lg, lfp, lgm, nx);
            for (int i = 1 - GHC; i <= nx + GHC; ++i) {
                AT(ssx2, i, j, k) = L(lfp, i); AT(ssx1, i, j, k) = L(lgm, i);
            }
        }
    /* y lines (i,:,k) */
    ORC_OMP("omp for collapse(2) schedule(static)")
    for (int k = ks; k <= ke; ++k)
        for (int i = 1; i <= nx; ++i) {
            for (int j = 1 - GHC; j <= ny + GHC; ++j) {
                L(lf, j) = AT(sy2, i, j, k); L(lg, j) = AT(sy1, i, j, k);
                L(lfp, j) = AT(ssy2, i, j, k); L(lgm, j) = AT(ssy1, i, j, k);
            }
            orc_crweno_flux_split(lf, lg, lfp, lgm, ny);
            for (int j = 1 - GHC; j <= ny + GHC; ++j) {
                AT(ssy2, i, j, k) = L(lfp, j); AT(ssy1, i, j, k) = L(lgm, j);
            }
        }
    /* z lines (i,j,:), 3D only */
    if (three_d) {
        ORC_OMP("omp for collapse(2) schedule(static)")
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                for (int k = ks - GHC; k <= ke + GHC; ++k) {
                    L(lf, k) = AT(sz2, i, j, k); L(lg, k) = AT(sz1, i, j, k);
                    L(lfp, k) = AT(ssz2, i, j, k); L(lgm, k) = AT(ssz1, i, j, k);
                }
                orc_crweno_flux_split(lf, lg, lfp, lgm, s->p.nz);
                for (int k = ks - GHC; k <= ke + GHC; ++k) {
                    AT(ssz2, i, j, k) = L(lfp, k); AT(ssz1, i, j, k) = L(lgm, k);
                }
            }
    }
    }
    free(scr);

    memcpy(sx1, ssx1, sizeof(double) * s->nbox);
    memcpy(sx2, ssx2, sizeof(double) * s->nbox);
    memcpy(sy1, ssy1, sizeof(double) * s->nbox);
    memcpy(sy2, ssy2, sizeof(double) * s->nbox);
    memcpy(sz1, ssz1, sizeof(double) * s->nbox);
    memcpy(sz2, ssz2, sizeof(double) * s->nbox);

    const double dx = s->p.dx, dy = s->p.dy, dz = s->p.dz;
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double r = (AT(sx1, i - 1, j, k) - AT(sx1, i, j, k)) / dx +
                           (AT(sx2, i - 1, j, k) - AT(sx2, i, j, k)) / dx +
                           (AT(sy1, i, j - 1, k) - AT(sy1, i, j, k)) / dy +
                           (AT(sy2, i, j - 1, k) - AT(sy2, i, j, k)) / dy;
                if (three_d)
                    r = r + (AT(sz1, i, j, k - 1) - AT(sz1, i, j, k)) / dz +
                        (AT(sz2, i, j, k - 1) - AT(sz2, i, j, k)) / dz;
                AT(tmp, i, j, k) = r;
            }
}

/* TVD-RK3, incremental form (level_set_rk3_solver.f90:1-94). */
static void rk3_rhs(orc_state *s)
{
    if (s->scheme == 1) orc_rk3_source_uccd(s);
    else orc_rk3_source(s);
}

void orc_rk3_advect(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const double dt = s->p.dt;
    double *phi = s->f[ORC_PHI], *tmp = s->f[ORC_PHI_TMP];
    double *l1 = s->f[ORC_L1], *l2 = s->f[ORC_L2], *l3 = s->f[ORC_L3];

    rk3_rhs(s);
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                AT(l1, i, j, k) = AT(tmp, i, j, k);
                double src = AT(l1, i, j, k);
                AT(phi, i, j, k) = AT(phi, i, j, k) + src * dt;
            }
    orc_bc(s, phi);

    rk3_rhs(s);
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                AT(l2, i, j, k) = AT(tmp, i, j, k);
                double src = (-3.0 * AT(l1, i, j, k) + AT(l2, i, j, k)) / 4.0;
                AT(phi, i, j, k) = AT(phi, i, j, k) + src * dt;
            }
    orc_bc(s, phi);

    rk3_rhs(s);
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                AT(l3, i, j, k) = AT(tmp, i, j, k);
                double src = (-AT(l1, i, j, k) - AT(l2, i, j, k) + 8.0 * AT(l3, i, j, k)) / 12.0;
                AT(phi, i, j, k) = AT(phi, i, j, k) + src * dt;
            }
    orc_bc(s, phi);
}

/* ------------------------------------------------------------------------
 * Reinitialisation (level_set_rk3_redis.f90).
 * ---------------------------------------------------------------------- */
double orc_phyn(double a, double b, double c, double d)
{
    const double eps = 1.0e-6;
    double is0 = 13.0 * sq(a - b) + 3.0 * sq(a - 3.0 * b);
    double is1 = 13.0 * sq(b - c) + 3.0 * sq(b + c);
    double is2 = 13.0 * sq(c - d) + 3.0 * sq(3.0 * c - d);
    double alp0 = 1.0 / sq(eps + is0);
    double alp1 = 6.0 / sq(eps + is1);
    double alp2 = 3.0 / sq(eps + is2);
    double w0 = alp0 / (alp0 + alp1 + alp2);
    double w2 = alp2 / (alp0 + alp1 + alp2);
    return w0 / 3.0 * (a - 2.0 * b + c) + (w2 - 0.5) / 6.0 * (b - 2.0 * c + d);
}

/* level_set_rk3_redis.f90:266-369 (3D: Vortex/ :357-359; 2D: Disk/ :357-361) */
void orc_redis_gradient(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const int three_d = (s->p.dim == 3);
    const double dx = s->p.dx, dy = s->p.dy, dz = s->p.dz;
    double *P = s->f[ORC_PHI];
    double *sx1 = s->f[ORC_SX1], *sx2 = s->f[ORC_SX2], *sy1 = s->f[ORC_SY1], *sy2 = s->f[ORC_SY2];
    double *sz1 = s->f[ORC_SZ1], *sz2 = s->f[ORC_SZ2];
    double *grad = s->f[ORC_GRAD], *S0 = s->f[ORC_SIGN_TMP];

    orc_bc(s, P);
#define PH(a, b, c) AT(P, (a), (b), (c))
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double a = 1.0 / (12.0 * dx) *
                           (-(PH(i - 1, j, k) - PH(i - 2, j, k)) + 7.0 * (PH(i, j, k) - PH(i - 1, j, k)) +
                            7.0 * (PH(i + 1, j, k) - PH(i, j, k)) - (PH(i + 2, j, k) - PH(i + 1, j, k)));
                double b = 1.0 / (12.0 * dy) *
                           (-(PH(i, j - 1, k) - PH(i, j - 2, k)) + 7.0 * (PH(i, j, k) - PH(i, j - 1, k)) +
                            7.0 * (PH(i, j + 1, k) - PH(i, j, k)) - (PH(i, j + 2, k) - PH(i, j + 1, k)));
                double c = 1.0 / (12.0 * dz) *
                           (-(PH(i, j, k - 1) - PH(i, j, k - 2)) + 7.0 * (PH(i, j, k) - PH(i, j, k - 1)) +
                            7.0 * (PH(i, j, k + 1) - PH(i, j, k)) - (PH(i, j, k + 2) - PH(i, j, k + 1)));

                double xs1 = a + 1.0 / dx *
                    orc_phyn(PH(i + 3, j, k) - 2.0 * PH(i + 2, j, k) + PH(i + 1, j, k),
                             PH(i + 2, j, k) - 2.0 * PH(i + 1, j, k) + PH(i, j, k),
                             PH(i + 1, j, k) - 2.0 * PH(i, j, k) + PH(i - 1, j, k),
                             PH(i, j, k) - 2.0 * PH(i - 1, j, k) + PH(i - 2, j, k));
                double xs2 = a - 1.0 / dx *
                    orc_phyn(PH(i - 3, j, k) - 2.0 * PH(i - 2, j, k) + PH(i - 1, j, k),
                             PH(i - 2, j, k) - 2.0 * PH(i - 1, j, k) + PH(i, j, k),
                             PH(i - 1, j, k) - 2.0 * PH(i, j, k) + PH(i + 1, j, k),
                             PH(i, j, k) - 2.0 * PH(i + 1, j, k) + PH(i + 2, j, k));
                double ys1 = b + 1.0 / dy *
                    orc_phyn(PH(i, j + 3, k) - 2.0 * PH(i, j + 2, k) + PH(i, j + 1, k),
                             PH(i, j + 2, k) - 2.0 * PH(i, j + 1, k) + PH(i, j, k),
                             PH(i, j + 1, k) - 2.0 * PH(i, j, k) + PH(i, j - 1, k),
                             PH(i, j, k) - 2.0 * PH(i, j - 1, k) + PH(i, j - 2, k));
                double ys2 = b - 1.0 / dy *
                    orc_phyn(PH(i, j - 3, k) - 2.0 * PH(i, j - 2, k) + PH(i, j - 1, k),
                             PH(i, j - 2, k) - 2.0 * PH(i, j - 1, k) + PH(i, j, k),
                             PH(i, j - 1, k) - 2.0 * PH(i, j, k) + PH(i, j + 1, k),
                             PH(i, j, k) - 2.0 * PH(i, j + 1, k) + PH(i, j + 2, k));
                double zs1 = c + 1.0 / dz *
                    orc_phyn(PH(i, j, k + 3) - 2.0 * PH(i, j, k + 2) + PH(i, j, k + 1),
                             PH(i, j, k + 2) - 2.0 * PH(i, j, k + 1) + PH(i, j, k),
                             PH(i, j, k + 1) - 2.0 * PH(i, j, k) + PH(i, j, k - 1),
                             PH(i, j, k) - 2.0 * PH(i, j, k - 1) + PH(i, j, k - 2));
                double zs2 = c - 1.0 / dz *
                    orc_phyn(PH(i, j, k - 3) - 2.0 * PH(i, j, k - 2) + PH(i, j, k - 1),
                             PH(i, j, k - 2) - 2.0 * PH(i, j, k - 1) + PH(i, j, k),
                             PH(i, j, k - 1) - 2.0 * PH(i, j, k) + PH(i, j, k + 1),
                             PH(i, j, k) - 2.0 * PH(i, j, k + 1) + PH(i, j, k + 2));
                AT(sx1, i, j, k) = xs1; AT(sx2, i, j, k) = xs2;
                AT(sy1, i, j, k) = ys1; AT(sy2, i, j, k) = ys2;
                AT(sz1, i, j, k) = zs1; AT(sz2, i, j, k) = zs2;

                double upm = -fmin(xs1, 0.0), upp = fmax(xs1, 0.0);
                double umm = -fmin(xs2, 0.0), ump = fmax(xs2, 0.0);
                double vpm = -fmin(ys1, 0.0), vpp = fmax(ys1, 0.0);
                double vmm = -fmin(ys2, 0.0), vmp = fmax(ys2, 0.0);
                double wpm = -fmin(zs1, 0.0), wpp = fmax(zs1, 0.0);
                double wmm = -fmin(zs2, 0.0), wmp = fmax(zs2, 0.0);
                double g;
                if (AT(S0, i, j, k) >= 0.0) {
                    if (three_d) g = sqrt(sq(fmax(upm, ump)) + sq(fmax(vpm, vmp)) + sq(fmax(wpm, wmp)));
                    else g = sqrt(sq(fmax(upm, ump)) + sq(fmax(vpm, vmp)));
                } else {
                    if (three_d) g = sqrt(sq(fmax(upp, umm)) + sq(fmax(vpp, vmm)) + sq(fmax(wpp, wmm)));
                    else g = sqrt(sq(fmax(upp, umm)) + sq(fmax(vpp, vmm)));
                }
                AT(grad, i, j, k) = g;
            }
#undef PH
}

/* level_set_rk3_redis.f90:371-454 (3D box weight 51: Vortex/ :432-435;
 * 2D weight 17 over kk = 0: Disk/ :434-438) */
void orc_redis_lambda(orc_state *s, int btn)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const int three_d = (s->p.dim == 3);
    double *s1 = s->f[ORC_SX1], *s2 = s->f[ORC_SX2], *s3 = s->f[ORC_S3];
    double *D = s->f[ORC_DELTA], *S0 = s->f[ORC_SIGN_TMP], *grad = s->f[ORC_GRAD];
    if (btn == 0) {
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) AT(s1, i, j, k) = 0.0;
        return;
    }
    orc_ls_funs(s);
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double lam = AT(D, i, j, k);
                AT(s2, i, j, k) = lam * AT(S0, i, j, k) * (AT(grad, i, j, k) - 1.0);
                AT(s3, i, j, k) = AT(grad, i, j, k) * AT(D, i, j, k) * lam;
            }
    orc_bc(s, s2);
    orc_bc(s, s3);
    const double wc = three_d ? 51.0 : 17.0;
    const int kr = three_d ? 1 : 0;
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double a = wc * AT(s2, i, j, k);
                double b = wc * AT(s3, i, j, k);
                for (int kk = -kr; kk <= kr; ++kk)
                    for (int jj = -1; jj <= 1; ++jj)
                        for (int ii = -1; ii <= 1; ++ii) {
                            a = a + AT(s2, i + ii, j + jj, k + kk);
                            b = b + AT(s3, i + ii, j + jj, k + kk);
                        }
                AT(s1, i, j, k) = 0.0;
                if (fabs(b) > 1.0e-12) AT(s1, i, j, k) = a / b * AT(D, i, j, k);
            }
}

/* level_set_rk3_redis.f90:176-264 */
static void orc_redis_solver(orc_state *s, int btn)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const double rdt = s->p.rdt;
    double *phi = s->f[ORC_PHI], *S0 = s->f[ORC_SIGN_TMP], *s1 = s->f[ORC_SX1];
    double *grad = s->f[ORC_GRAD];
    double *l1 = s->f[ORC_L1], *l2 = s->f[ORC_L2], *l3 = s->f[ORC_L3];

    orc_redis_gradient(s);
    orc_redis_lambda(s, btn);
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                AT(l1, i, j, k) = (AT(S0, i, j, k) - AT(s1, i, j, k)) * AT(grad, i, j, k) - AT(S0, i, j, k);
                double src = AT(l1, i, j, k);
                AT(phi, i, j, k) = AT(phi, i, j, k) - rdt * src;
            }
    orc_bc(s, phi);

    orc_redis_gradient(s);
    orc_redis_lambda(s, btn);
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                AT(l2, i, j, k) = (AT(S0, i, j, k) - AT(s1, i, j, k)) * AT(grad, i, j, k) - AT(S0, i, j, k);
                double src = (-3.0 * AT(l1, i, j, k) + AT(l2, i, j, k)) / 4.0;
                AT(phi, i, j, k) = AT(phi, i, j, k) - rdt * src;
            }
    orc_bc(s, phi);

    orc_redis_gradient(s);
    orc_redis_lambda(s, btn);
    ORC_PAR
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                AT(l3, i, j, k) = (AT(S0, i, j, k) - AT(s1, i, j, k)) * AT(grad, i, j, k) - AT(S0, i, j, k);
                double src = (-AT(l1, i, j, k) - AT(l2, i, j, k) + 8.0 * AT(l3, i, j, k)) / 12.0;
                AT(phi, i, j, k) = AT(phi, i, j, k) - rdt * src;
            }
    orc_bc(s, phi);
}

int orc_redis(orc_state *s, int btn, double T)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    double *phi = s->f[ORC_PHI], *tmp = s->f[ORC_PHI_TMP];
    double *S0 = s->f[ORC_SIGN_TMP], *SG = s->f[ORC_SIGN], *grad = s->f[ORC_GRAD];

    /* level_set_redis_init (:81-114) */
    orc_bc(s, phi);
    orc_ls_funs(s);
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) AT(S0, i, j, k) = AT(SG, i, j, k);
    if (btn == 0) {
        /* level_set_redis_stable (:116-174) */
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) AT(S0, i, j, k) = AT(phi, i, j, k);
        orc_redis_gradient(s);
        double gmax = 0.0;
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) gmax = fmax(gmax, AT(grad, i, j, k));
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) AT(phi, i, j, k) = AT(phi, i, j, k) / gmax;
        orc_bc(s, phi);
    }

    double timestop;
    if (btn == 0) {
        double ex = s->p.xend - s->p.xstart, ey = s->p.yend - s->p.ystart, ez = s->p.zend - s->p.zstart;
        double m = ex; if (ey > m) m = ey; if (ez > m) m = ez;
        timestop = 1.5 * m;
    } else {
        timestop = 2.0 * s->p.ls_wid;
    }
    if (T >= 0.0) timestop = T;

    int iter = 0;
    double time = 0.0;
    for (;;) {
        iter = iter + 1;
        time = time + s->p.rdt;
        ORC_PAR
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) AT(tmp, i, j, k) = AT(phi, i, j, k);
        orc_redis_solver(s, btn);
        double error = 0.0;
        ORC_OMP("omp parallel for collapse(2) reduction(max:error) schedule(static)")
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i)
                    error = fmax(error, fabs(AT(tmp, i, j, k) - AT(phi, i, j, k)));
        orc_bc(s, phi);
        s->red_error = error;
        if (time > timestop || error <= 1.0e-6) break;
    }
    s->last_redis_iters = iter;
    return iter;
}

/* ------------------------------------------------------------------------
 * Combined compact difference (CCD) line solver.
 * ---------------------------------------------------------------------- */
typedef struct {
    int n;
    double dx;
    double *A, *B, *AA, *BB;   /* (3, n) column-major: X[m + 3*i], m = 0..2 */
} ccd_mat;

#define M3(X, m, i) (X)[(m) + 3 * (i)]

/* matrix_solver.f90:6-45 */
static void twin_dec(double *A, double *B, double *AA, double *BB, int n)
{
    const int IS = 0, IE = n - 1;
    M3(A, 0, IS) = 0.0; M3(AA, 0, IS) = 0.0; M3(A, 2, IE) = 0.0; M3(AA, 2, IE) = 0.0;
    M3(B, 0, IS) = 0.0; M3(BB, 0, IS) = 0.0; M3(B, 2, IE) = 0.0; M3(BB, 2, IE) = 0.0;
    for (int i = IS + 1; i <= IE; ++i) {
        double Den = M3(A, 1, i - 1) * M3(BB, 1, i - 1) - M3(AA, 1, i - 1) * M3(B, 1, i - 1);
        double SC1 = -(M3(AA, 1, i - 1) * M3(B, 0, i) - M3(A, 0, i) * M3(BB, 1, i - 1)) / Den;
        double SC2 = (M3(B, 0, i) * M3(A, 1, i - 1) - M3(B, 1, i - 1) * M3(A, 0, i)) / Den;
        M3(A, 0, i) = SC1;
        M3(A, 1, i) = M3(A, 1, i) - (SC1 * M3(A, 2, i - 1) + SC2 * M3(AA, 2, i - 1));
        M3(B, 0, i) = SC2;
        M3(B, 1, i) = M3(B, 1, i) - (SC1 * M3(B, 2, i - 1) + SC2 * M3(BB, 2, i - 1));
        SC1 = -(M3(AA, 1, i - 1) * M3(BB, 0, i) - M3(AA, 0, i) * M3(BB, 1, i - 1)) / Den;
        SC2 = (M3(BB, 0, i) * M3(A, 1, i - 1) - M3(B, 1, i - 1) * M3(AA, 0, i)) / Den;
        M3(AA, 0, i) = SC1;
        M3(AA, 1, i) = M3(AA, 1, i) - (SC1 * M3(A, 2, i - 1) + SC2 * M3(AA, 2, i - 1));
        M3(BB, 0, i) = SC2;
        M3(BB, 1, i) = M3(BB, 1, i) - (SC1 * M3(B, 2, i - 1) + SC2 * M3(BB, 2, i - 1));
    }
}

/* matrix_solver.f90:47-78 */
static void twin_bks(const double *A, const double *B, const double *AA, const double *BB,
                     double *S, double *SS, int n)
{
    const int IS = 0, IE = n - 1;
    for (int i = IS + 1; i <= IE; ++i) {
        S[i] = S[i] - (M3(A, 0, i) * S[i - 1] + M3(B, 0, i) * SS[i - 1]);
        SS[i] = SS[i] - (M3(AA, 0, i) * S[i - 1] + M3(BB, 0, i) * SS[i - 1]);
    }
    double Den = M3(A, 1, IE) * M3(BB, 1, IE) - M3(AA, 1, IE) * M3(B, 1, IE);
    double SolS = -(M3(B, 1, IE) * SS[IE] - M3(BB, 1, IE) * S[IE]) / Den;
    double SolSS = (M3(A, 1, IE) * SS[IE] - M3(AA, 1, IE) * S[IE]) / Den;
    S[IE] = SolS;
    SS[IE] = SolSS;
    for (int i = IE - 1; i >= IS; --i) {
        S[i] = S[i] - (M3(A, 2, i) * SolS + M3(B, 2, i) * SolSS);
        SS[i] = SS[i] - (M3(AA, 2, i) * SolS + M3(BB, 2, i) * SolSS);
        Den = M3(A, 1, i) * M3(BB, 1, i) - M3(AA, 1, i) * M3(B, 1, i);
        SolS = -(M3(B, 1, i) * SS[i] - M3(BB, 1, i) * S[i]) / Den;
        SolSS = (M3(A, 1, i) * SS[i] - M3(AA, 1, i) * S[i]) / Den;
        S[i] = SolS;
        SS[i] = SolSS;
    }
}

/* ccd_solvers.f90:80-134 (bc) and 136-163 (init_ccd), then TWIN_DEC once
 * (root_init, :382-392) */
static void ccd_build(ccd_mat *m, int n, double dx)
{
    m->n = n; m->dx = dx;
    m->A = (double *)calloc((size_t)(3 * n), sizeof(double));
    m->B = (double *)calloc((size_t)(3 * n), sizeof(double));
    m->AA = (double *)calloc((size_t)(3 * n), sizeof(double));
    m->BB = (double *)calloc((size_t)(3 * n), sizeof(double));
    const int is = 0, ie = n - 1;
    M3(m->A, 1, is) = 1.0;  M3(m->B, 1, is) = 0.0;
    M3(m->AA, 1, is) = 0.0; M3(m->BB, 1, is) = 1.0;
    M3(m->A, 2, is) = 2.0;  M3(m->B, 2, is) = -dx;
    M3(m->AA, 2, is) = -2.5 / dx; M3(m->BB, 2, is) = 8.5;
    M3(m->A, 1, ie) = 1.0;  M3(m->B, 1, ie) = 0.0;
    M3(m->AA, 1, ie) = 0.0; M3(m->BB, 1, ie) = 1.0;
    M3(m->A, 0, ie) = 2.0;  M3(m->B, 0, ie) = dx;
    M3(m->AA, 0, ie) = 2.5 / dx; M3(m->BB, 0, ie) = 8.5;
    for (int i = is + 1; i <= ie - 1; ++i) {
        M3(m->A, 0, i) = 7.0 / 16.0; M3(m->A, 1, i) = 1.0; M3(m->A, 2, i) = 7.0 / 16.0;
        M3(m->B, 0, i) = dx / 16.0;  M3(m->B, 1, i) = 0.0; M3(m->B, 2, i) = -dx / 16.0;
        M3(m->AA, 0, i) = -(9.0 / 8.0 / dx); M3(m->AA, 1, i) = 0.0; M3(m->AA, 2, i) = 9.0 / 8.0 / dx;
        M3(m->BB, 0, i) = -1.0 / 8.0; M3(m->BB, 1, i) = 1.0; M3(m->BB, 2, i) = -1.0 / 8.0;
    }
    twin_dec(m->A, m->B, m->AA, m->BB, n);
}

static void ccd_free(ccd_mat *m)
{
    free(m->A); free(m->B); free(m->AA); free(m->BB);
}

/* asgn_src_ccd (ccd_solvers.f90:291-310) + TWIN_BKS (root_solve :421-427) */
static void ccd_solve(const ccd_mat *m, const double *f, double *fx, double *fxx)
{
    const int n = m->n, is = 0, ie = n - 1;
    const double dx = m->dx;
    double *S = fx, *SS = fxx;   /* su/ssu are copied out unchanged */
    S[is] = (-3.5 * f[is] + 4.0 * f[is + 1] - 0.5 * f[is + 2]) / dx;
    SS[is] = (34.0 / 3.0 * f[is] - 83.0 / 4.0 * f[is + 1] + 10.0 * f[is + 2] - 7.0 / 12.0 * f[is + 3]) /
             (dx * dx);
    for (int i = is + 1; i <= ie - 1; ++i) {
        S[i] = 15.0 / 16.0 * (-f[i - 1] + f[i + 1]) / dx;
        SS[i] = (3.0 * f[i - 1] - 6.0 * f[i] + 3.0 * f[i + 1]) / (dx * dx);
    }
    S[ie] = -(-3.5 * f[ie] + 4.0 * f[ie - 1] - 0.5 * f[ie - 2]) / dx;
    SS[ie] = (34.0 / 3.0 * f[ie] - 83.0 / 4.0 * f[ie - 1] + 10.0 * f[ie - 2] - 7.0 / 12.0 * f[ie - 3]) /
             (dx * dx);
    twin_bks(m->A, m->B, m->AA, m->BB, S, SS, n);
}

void orc_ccd_line(int n, double dx, const double *f, double *fx, double *fxx)
{
    ccd_mat m;
    ccd_build(&m, n, dx);
    ccd_solve(&m, f, fx, fxx);
    ccd_free(&m);
}

/* Upwind CCD (ccd_solvers.f90:165-214 init_uccd, :80-134 bc, :312-345
 * asgn_src_uccd, root_solve "uccd" :428-438): two block systems with the CCD
 * boundary rows, each factorised once by TWIN_DEC; fx(i) takes the upwind
 * system's value for u(i) > 0 and the downwind one's otherwise (fxx is not
 * returned: nothing on the advection path reads it). */
static void uccd_build(ccd_mat *mu, ccd_mat *md, int n, double dx)
{
    const double A1 = 0.875, B1 = 0.1251282341599089, B2 = -0.2487176584009104, B3 = 0.0001282341599089;
    ccd_mat *ms[2] = {mu, md};
    for (int w = 0; w < 2; ++w) {
        ccd_mat *m = ms[w];
        m->n = n; m->dx = dx;
        m->A = (double *)calloc((size_t)(3 * n), sizeof(double));
        m->B = (double *)calloc((size_t)(3 * n), sizeof(double));
        m->AA = (double *)calloc((size_t)(3 * n), sizeof(double));
        m->BB = (double *)calloc((size_t)(3 * n), sizeof(double));
        const int is = 0, ie = n - 1;
        M3(m->A, 1, is) = 1.0;  M3(m->B, 1, is) = 0.0;
        M3(m->AA, 1, is) = 0.0; M3(m->BB, 1, is) = 1.0;
        M3(m->A, 2, is) = 2.0;  M3(m->B, 2, is) = -dx;
        M3(m->AA, 2, is) = -2.5 / dx; M3(m->BB, 2, is) = 8.5;
        M3(m->A, 1, ie) = 1.0;  M3(m->B, 1, ie) = 0.0;
        M3(m->AA, 1, ie) = 0.0; M3(m->BB, 1, ie) = 1.0;
        M3(m->A, 0, ie) = 2.0;  M3(m->B, 0, ie) = dx;
        M3(m->AA, 0, ie) = 2.5 / dx; M3(m->BB, 0, ie) = 8.5;
        for (int i = is + 1; i <= ie - 1; ++i) {
            if (w == 0) {
                M3(m->A, 0, i) = A1;       M3(m->A, 1, i) = 1.0;      M3(m->A, 2, i) = 0.0;
                M3(m->B, 0, i) = B1 * dx;  M3(m->B, 1, i) = B2 * dx;  M3(m->B, 2, i) = B3 * dx;
            } else {
                M3(m->A, 0, i) = 0.0;      M3(m->A, 1, i) = 1.0;      M3(m->A, 2, i) = A1;
                M3(m->B, 0, i) = -B3 * dx; M3(m->B, 1, i) = -B2 * dx; M3(m->B, 2, i) = -B1 * dx;
            }
            M3(m->AA, 0, i) = -(9.0 / 8.0 / dx); M3(m->AA, 1, i) = 0.0; M3(m->AA, 2, i) = 9.0 / 8.0 / dx;
            M3(m->BB, 0, i) = -1.0 / 8.0; M3(m->BB, 1, i) = 1.0; M3(m->BB, 2, i) = -1.0 / 8.0;
        }
        twin_dec(m->A, m->B, m->AA, m->BB, n);
    }
}

static void uccd_solve(const ccd_mat *mu, const ccd_mat *md, const double *f, const double *u, double *fx,
                       double *work)
{
    const int n = mu->n, is = 0, ie = n - 1;
    const double dx = mu->dx;
    const double c3 = -0.06096119008109, c2 = 1.99692238016218, c1 = -1.93596119008109;
    double *SU = work, *SSU = work + n, *SD = work + 2 * n, *SSD = work + 3 * n;
    SU[is] = (-3.5 * f[is] + 4.0 * f[is + 1] - 0.5 * f[is + 2]) / dx;
    SSU[is] = (34.0 / 3.0 * f[is] - 83.0 / 4.0 * f[is + 1] + 10.0 * f[is + 2] - 7.0 / 12.0 * f[is + 3]) /
              (dx * dx);
    SD[is] = SU[is];
    SSD[is] = SSU[is];
    for (int i = is + 1; i <= ie - 1; ++i) {
        SU[i] = (c1 * f[i - 1] + c2 * f[i] + c3 * f[i + 1]) / dx;
        SSU[i] = (3.0 * f[i - 1] - 6.0 * f[i] + 3.0 * f[i + 1]) / (dx * dx);
        SD[i] = -(c3 * f[i - 1] + c2 * f[i] + c1 * f[i + 1]) / dx;
        SSD[i] = SSU[i];
    }
    SU[ie] = -(-3.5 * f[ie] + 4.0 * f[ie - 1] - 0.5 * f[ie - 2]) / dx;
    SSU[ie] = (34.0 / 3.0 * f[ie] - 83.0 / 4.0 * f[ie - 1] + 10.0 * f[ie - 2] - 7.0 / 12.0 * f[ie - 3]) /
              (dx * dx);
    SD[ie] = SU[ie];
    SSD[ie] = SSU[ie];
    twin_bks(mu->A, mu->B, mu->AA, mu->BB, SU, SSU, n);
    twin_bks(md->A, md->B, md->AA, md->BB, SD, SSD, n);
    for (int i = is; i <= ie; ++i) fx[i] = u[i] > 0.0 ? SU[i] : SD[i];
}

void orc_uccd_line(int n, double dx, const double *f, const double *u, double *fx)
{
    ccd_mat mu, md;
    uccd_build(&mu, &md, n, dx);
    double *work = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    uccd_solve(&mu, &md, f, u, fx, work);
    free(work);
    ccd_free(&mu);
    ccd_free(&md);
}

/* level_set_rk3_source_ccd (level_set_rk3_solver.f90:223-286): UCCD first
 * derivatives along every interior x, y (and, in 3D, z) line of the box, then
 * tmp = -u phi_x - v phi_y - w phi_z.  In 2D the reference also solves the
 * 7-point z lines, whose term is multiplied by w = 0; it is omitted here. */
void orc_rk3_source_uccd(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const int three_d = (s->p.dim == 3);
    double *phi = s->f[ORC_PHI], *tmp = s->f[ORC_PHI_TMP];
    double *U = s->f[ORC_U], *V = s->f[ORC_V], *W = s->f[ORC_W];
    double *sx = s->f[ORC_SX1], *sy = s->f[ORC_SY1], *sz = s->f[ORC_SZ1];
    const int nmax = (s->nxb > s->nyb ? (s->nxb > s->nzb ? s->nxb : s->nzb) : (s->nyb > s->nzb ? s->nyb : s->nzb));
    double *f = (double *)malloc(sizeof(double) * nmax), *u = (double *)malloc(sizeof(double) * nmax);
    double *fx = (double *)malloc(sizeof(double) * nmax), *work = (double *)malloc(sizeof(double) * 4 * nmax);
    ccd_mat mu, md;
    /* all three solvers are built with dx (ccd_solvers.f90:406-408) */
    uccd_build(&mu, &md, s->nxb, s->p.dx);
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j) {
            for (int i = 1 - GHC; i <= nx + GHC; ++i) { f[i - 1 + GHC] = AT(phi, i, j, k); u[i - 1 + GHC] = AT(U, i, j, k); }
            uccd_solve(&mu, &md, f, u, fx, work);
            for (int i = 1 - GHC; i <= nx + GHC; ++i) AT(sx, i, j, k) = fx[i - 1 + GHC];
        }
    ccd_free(&mu); ccd_free(&md);
    uccd_build(&mu, &md, s->nyb, s->p.dx);
    for (int k = ks; k <= ke; ++k)
        for (int i = 1; i <= nx; ++i) {
            for (int j = 1 - GHC; j <= ny + GHC; ++j) { f[j - 1 + GHC] = AT(phi, i, j, k); u[j - 1 + GHC] = AT(V, i, j, k); }
            uccd_solve(&mu, &md, f, u, fx, work);
            for (int j = 1 - GHC; j <= ny + GHC; ++j) AT(sy, i, j, k) = fx[j - 1 + GHC];
        }
    ccd_free(&mu); ccd_free(&md);
    if (three_d) {
        uccd_build(&mu, &md, s->nzb, s->p.dx);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                for (int k = 1 - GHC; k <= s->p.nz + GHC; ++k) { f[k - 1 + GHC] = AT(phi, i, j, k); u[k - 1 + GHC] = AT(W, i, j, k); }
                uccd_solve(&mu, &md, f, u, fx, work);
                for (int k = 1 - GHC; k <= s->p.nz + GHC; ++k) AT(sz, i, j, k) = fx[k - 1 + GHC];
            }
        ccd_free(&mu); ccd_free(&md);
    }
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double r = -AT(U, i, j, k) * AT(sx, i, j, k) - AT(V, i, j, k) * AT(sy, i, j, k);
                if (three_d) r = r - AT(W, i, j, k) * AT(sz, i, j, k);
                AT(tmp, i, j, k) = r;
            }
    free(f); free(u); free(fx); free(work);
}

/* surface_norms: first and second CCD derivatives along every x, y and z line
 * of phi including ghost lines, then bc on the results (derivatives.f90:1-50,
 * 96-116).  The three cross-derivative sweeps (:52-94) write only xy, xz, yz
 * and curv, which nothing on the MPLS path reads, so they are not restated.
 * All three solvers use dx (ccd_solvers.f90:406-408). */
void orc_surface_norms(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const double dx = s->p.dx;
    double *P = s->f[ORC_PHI];
    double *NX = s->f[ORC_NX], *NY = s->f[ORC_NY], *NZ = s->f[ORC_NZ];
    double *NXX = s->f[ORC_NXX], *NYY = s->f[ORC_NYY], *NZZ = s->f[ORC_NZZ];
    ccd_mat mx, my, mz;
    ccd_build(&mx, s->nxb, dx);
    ccd_build(&my, s->nyb, dx);
    ccd_build(&mz, s->nzb, dx);
    int len = s->nxb; if (s->nyb > len) len = s->nyb; if (s->nzb > len) len = s->nzb;
    double *scr = (double *)malloc(sizeof(double) * 3 * len * orc_threads());
    ORC_OMP("omp parallel")
    {
    double *f = scr + 3 * len * orc_tid(), *fx = f + len, *fxx = fx + len;

    ORC_OMP("omp for collapse(2) schedule(static)")
    for (int k = ks - GHC; k <= ke + GHC; ++k)
        for (int j = 1 - GHC; j <= ny + GHC; ++j) {
            for (int i = 1 - GHC; i <= nx + GHC; ++i) f[i - 1 + GHC] = AT(P, i, j, k);
            ccd_solve(&mx, f, fx, fxx);
            for (int i = 1 - GHC; i <= nx + GHC; ++i) {
                AT(NX, i, j, k) = fx[i - 1 + GHC]; AT(NXX, i, j, k) = fxx[i - 1 + GHC];
            }
        }
    ORC_OMP("omp for collapse(2) schedule(static)")
    for (int k = ks - GHC; k <= ke + GHC; ++k)
        for (int i = 1 - GHC; i <= nx + GHC; ++i) {
            for (int j = 1 - GHC; j <= ny + GHC; ++j) f[j - 1 + GHC] = AT(P, i, j, k);
            ccd_solve(&my, f, fx, fxx);
            for (int j = 1 - GHC; j <= ny + GHC; ++j) {
                AT(NY, i, j, k) = fx[j - 1 + GHC]; AT(NYY, i, j, k) = fxx[j - 1 + GHC];
            }
        }
    ORC_OMP("omp for collapse(2) schedule(static)")
    for (int j = 1 - GHC; j <= ny + GHC; ++j)
        for (int i = 1 - GHC; i <= nx + GHC; ++i) {
            for (int k = ks - GHC; k <= ke + GHC; ++k) f[k - ks + GHC] = AT(P, i, j, k);
            ccd_solve(&mz, f, fx, fxx);
            for (int k = ks - GHC; k <= ke + GHC; ++k) {
                AT(NZ, i, j, k) = fx[k - ks + GHC]; AT(NZZ, i, j, k) = fxx[k - ks + GHC];
            }
        }
    }
    free(scr);
    ccd_free(&mx); ccd_free(&my); ccd_free(&mz);
    orc_bc(s, NX); orc_bc(s, NXX);
    orc_bc(s, NY); orc_bc(s, NYY);
    orc_bc(s, NZ); orc_bc(s, NZZ);
}

/* mass_preserving_level_set.f90:1-74 (2D |grad| omits z: Disk/ :24-26) */
int orc_mass_fix(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const int three_d = (s->p.dim == 3);
    const double r12 = s->p.rho_12, dx = s->p.dx, dy = s->p.dy, dz = s->p.dz;
    double *phi = s->f[ORC_PHI], *grad = s->f[ORC_GRAD], *D = s->f[ORC_DELTA], *H = s->f[ORC_HEAVY];
    double *NX = s->f[ORC_NX], *NY = s->f[ORC_NY], *NZ = s->f[ORC_NZ];
    int iter;
    for (iter = 1; iter <= 15; ++iter) {
        orc_surface_norms(s);
        orc_ls_mv(s);
        double lam = 0.0;
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double g;
                    if (three_d) g = sqrt(sq(AT(NX, i, j, k)) + sq(AT(NY, i, j, k)) + sq(AT(NZ, i, j, k)));
                    else g = sqrt(sq(AT(NX, i, j, k)) + sq(AT(NY, i, j, k)));
                    AT(grad, i, j, k) = g;
                    double plam = sq(AT(D, i, j, k)) * g;
                    if (s->p.inverse)
                        plam = plam * (2.0 * (r12 - 1.0) * AT(H, i, j, k) + 1.0) * dx * dy * dz;
                    else
                        plam = plam * (2.0 * (1.0 - r12) * AT(H, i, j, k) + r12) * dx * dy * dz;
                    lam = lam + plam;
                }
        lam = (s->imass - s->mass) / lam;
        ORC_PAR
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i)
                    AT(phi, i, j, k) = AT(phi, i, j, k) + lam * AT(D, i, j, k) * AT(grad, i, j, k);
        orc_bc(s, phi);
        if (fabs(s->imass - s->mass) / s->imass < 1.0e-14) break;
    }
    s->last_mass_iters = iter > 15 ? 15 : iter;
    return s->last_mass_iters;
}

/* output.f90:233-248 */
static double heavyside(double x)
{
    const double eps = 1.0e-12, pi = acos(-1.0);
    if (x > 1.0 - eps) return 1.0;
    if (x < -1.0 + eps) return 0.0;
    return 0.5 * ((1.0 + x) + sin(pi * x) / pi);
}

/* output.f90:169-204 and :226 (accumulator starts at 0; the reference leaves
 * it uninitialised, output.f90:172-178) */
double orc_interface_l1(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    double *phi = s->f[ORC_PHI], *p0 = s->f[ORC_PHI0];
    orc_ls_mv(s);
    double error = 0.0;
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double src = heavyside(AT(p0, i, j, k) / s->p.ls_wid);
                double tgt = heavyside(AT(phi, i, j, k) / s->p.ls_wid);
                error = error + fabs(src - tgt);
            }
    return error / (double)nx / (double)ny / (double)s->p.nz;
}

/* ------------------------------------------------------------------------
 * energy_and_momentum (Vortex/src/others/output.f90:250-397; the Disk/ and
 * Vortex2D/ copies are the same text): rho_mu (ns_solver/others.f90:47-80)
 * on the marker, the nine first CCD derivatives of the velocity along the
 * interior lines (:262-300; all three solvers use dx, ccd_solvers.f90:406-408),
 * then one pass over the interior cells (:313-360).  out[0] A (already / dx,
 * :364), out[1] v, out[2..4] px / v, py / v, pz / v, out[5..7] xc / v, yc / v,
 * zc / v, out[8] Ke, out[9] Po, out[10] Q -- the integrals before the deck's
 * Fr, We, Re and energy_unit scaling (:374-386).  The reference adds to px,
 * py, pz without clearing them (an uninitialised local); they start at 0
 * here.  marker = heavy (ls_funs of phi) or, with vof_marker, the VOF
 * (method 3).  Cell coordinates are coord()'s.
 * ---------------------------------------------------------------------- */
static int orc_cmp_int(int a, int b) { return a > b ? a : b; }
static double coord(double start, double end, double d, int n, int i);

void orc_energy_momentum(orc_state *s, double mu_12, int vof_marker, double *out)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const double dx = s->p.dx;
    const double dv = s->p.dx * s->p.dy * s->p.dz;
    double *P = s->f[ORC_PHI], *H = s->f[ORC_HEAVY], *VF = s->f[ORC_VOF];
    double *vel[3] = {s->f[ORC_U], s->f[ORC_V], s->f[ORC_W]};
    orc_ls_funs(s);   /* rho_mu calls ls_funs first (others.f90:53) */
    /* ten[3 f + d]: d/dx_d of velocity component f on the box */
    double *ten[9];
    for (int q = 0; q < 9; ++q) ten[q] = (double *)calloc((size_t)s->nbox, sizeof(double));
    ccd_mat m[3];
    ccd_build(&m[0], s->nxb, dx);
    ccd_build(&m[1], s->nyb, dx);
    ccd_build(&m[2], s->nzb, dx);
    const int len = orc_cmp_int(orc_cmp_int(s->nxb, s->nyb), s->nzb);
    double *f = (double *)malloc(sizeof(double) * 3 * len), *fx = f + len, *fxx = fx + len;
    for (int c = 0; c < 3; ++c) {
        double *U = vel[c];
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j) {
                for (int i = 1 - GHC; i <= nx + GHC; ++i) f[i - 1 + GHC] = AT(U, i, j, k);
                ccd_solve(&m[0], f, fx, fxx);
                for (int i = 1 - GHC; i <= nx + GHC; ++i) AT(ten[3 * c + 0], i, j, k) = fx[i - 1 + GHC];
            }
        for (int k = ks; k <= ke; ++k)
            for (int i = 1; i <= nx; ++i) {
                for (int j = 1 - GHC; j <= ny + GHC; ++j) f[j - 1 + GHC] = AT(U, i, j, k);
                ccd_solve(&m[1], f, fx, fxx);
                for (int j = 1 - GHC; j <= ny + GHC; ++j) AT(ten[3 * c + 1], i, j, k) = fx[j - 1 + GHC];
            }
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                for (int k = ks - GHC; k <= ke + GHC; ++k) f[k - ks + GHC] = AT(U, i, j, k);
                ccd_solve(&m[2], f, fx, fxx);
                for (int k = ks - GHC; k <= ke + GHC; ++k) AT(ten[3 * c + 2], i, j, k) = fx[k - ks + GHC];
            }
    }
    free(f);
    ccd_free(&m[0]); ccd_free(&m[1]); ccd_free(&m[2]);
    double A = 0.0, Ke = 0.0, Po = 0.0, Q = 0.0, v = 0.0, xc = 0.0, yc = 0.0, zc = 0.0;
    double px = 0.0, py = 0.0, pz = 0.0;
    const double r12 = s->p.rho_12;
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double phi = AT(P, i, j, k);
                if (phi < dx && phi > 0.0) A = A + dv;
                const double marker = vof_marker ? AT(VF, i, j, k) : AT(H, i, j, k);
                const double u = AT(vel[0], i, j, k), vv = AT(vel[1], i, j, k), w = AT(vel[2], i, j, k);
                const double x = coord(s->p.xstart, s->p.xend, s->p.dx, nx, i);
                const double y = coord(s->p.ystart, s->p.yend, s->p.dy, ny, j);
                const double z = coord(s->p.zstart, s->p.zend, s->p.dz, s->p.nz, k);
                px = px + marker * u * dv;
                py = py + marker * vv * dv;
                pz = pz + marker * w * dv;
                v = v + marker * dv;
                xc = xc + marker * x * dv;
                yc = yc + marker * y * dv;
                zc = zc + marker * z * dv;
                if (marker > 0.99) {
                    double rho, mu;
                    if (!s->p.inverse) {
                        rho = marker + r12 * (1.0 - marker);
                        mu = marker + mu_12 * (1.0 - marker);
                    } else {
                        rho = marker * r12 + (1.0 - marker);
                        mu = marker * mu_12 + (1.0 - marker);
                    }
                    const double lke = u * u + vv * vv + w * w;
                    Ke = Ke + lke * rho * 0.5 * dv;
                    Po = Po + z * rho * dv;
                    const double ux = AT(ten[0], i, j, k), uy = AT(ten[1], i, j, k), uz = AT(ten[2], i, j, k);
                    const double vx = AT(ten[3], i, j, k), vy = AT(ten[4], i, j, k), vz = AT(ten[5], i, j, k);
                    const double wx = AT(ten[6], i, j, k), wy = AT(ten[7], i, j, k), wz = AT(ten[8], i, j, k);
                    const double lq = 0.5 * (sq(uy + vx) + sq(uz + wx) + sq(wy + vz)) + sq(ux) + sq(vy) + sq(wz);
                    Q = Q + lq * 2.0 * mu * dv;
                }
            }
    for (int q = 0; q < 9; ++q) free(ten[q]);
    out[0] = A / dx;
    out[1] = v;
    out[2] = px / v; out[3] = py / v; out[4] = pz / v;
    out[5] = xc / v; out[6] = yc / v; out[7] = zc / v;
    out[8] = Ke; out[9] = Po; out[10] = Q;
}

/* ------------------------------------------------------------------------
 * Inputs.
 * ---------------------------------------------------------------------- */
/* cell-centre coordinate x(i) = xstart + (i - 0.5)*dx with (i - 0.5) in single
 * precision (Vortex/lib/data/tree.f90:254-274); index 0 and n+1 are the walls */
static double coord(double start, double end, double d, int n, int i)
{
    if (i <= 0) return start;
    if (i >= n + 1) return end;
    float h = (float)i - 0.5f;
    return start + (double)h * d;
}

void orc_velocity(orc_state *s, int field, double time, double t2s)
{
    const double pi = acos(-1.0);
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz, ks = s->ks, ke = s->ke;
    double *u = s->f[ORC_U], *v = s->f[ORC_V], *w = s->f[ORC_W];
    if (field == 0) {
        /* Disk/src/setup/velocity.f90:63-114 */
        double omega = 2.0 * pi / t2s;
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double x = coord(s->p.xstart, s->p.xend, s->p.dx, nx, i);
                    double y = coord(s->p.ystart, s->p.yend, s->p.dy, ny, j);
                    AT(u, i, j, k) = -omega * (y - 0.5);
                    AT(v, i, j, k) = omega * (x - 0.5);
                }
    } else {
        double ct = cos(pi * time / t2s);
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double x = coord(s->p.xstart, s->p.xend, s->p.dx, nx, i);
                    double y = coord(s->p.ystart, s->p.yend, s->p.dy, ny, j);
                    if (field == 1) {
                        /* Vortex2D/src/setup/velocity.f90:35-36 */
                        AT(u, i, j, k) = -sq(sin(pi * x)) * sin(2.0 * pi * y) * ct;
                        AT(v, i, j, k) = sq(sin(pi * y)) * sin(2.0 * pi * x) * ct;
                    } else {
                        /* Vortex/src/setup/velocity.f90:27-29 */
                        double z = coord(s->p.zstart, s->p.zend, s->p.dz, nz, k);
                        AT(u, i, j, k) = 2.0 * sq(sin(pi * x)) * sin(2.0 * pi * y) * sin(2.0 * pi * z) * ct;
                        AT(v, i, j, k) = -sin(2.0 * pi * x) * sq(sin(pi * y)) * sin(2.0 * pi * z) * ct;
                        AT(w, i, j, k) = -sin(2.0 * pi * x) * sin(2.0 * pi * y) * sq(sin(pi * z)) * ct;
                    }
                }
    }
    orc_bc(s, u); orc_bc(s, v); orc_bc(s, w);
}

void orc_initial_phi(orc_state *s, int shape, int with_vof)
{
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz, ks = s->ks, ke = s->ke;
    const double dx = s->p.dx, dy = s->p.dy, dz = s->p.dz;
    const double c015 = (double)0.15f, c035 = (double)0.35f, c0025 = (double)0.025f;
    const double c08 = (double)0.8f, r2 = (double)(0.15f * 0.15f);
    double *phi = s->f[ORC_PHI], *vof = s->f[ORC_VOF];
    const int ug = 30;
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double xc = coord(s->p.xstart, s->p.xend, dx, nx, i);
                double yc = coord(s->p.ystart, s->p.yend, dy, ny, j);
                double zc = coord(s->p.zstart, s->p.zend, dz, nz, k);
                if (with_vof || shape == 0) {
                    /* problem_init.f90:41-59: 30^3 sub-samples */
                    double xm = coord(s->p.xstart, s->p.xend, dx, nx, i - 1);
                    double ym = coord(s->p.ystart, s->p.yend, dy, ny, j - 1);
                    double zm = coord(s->p.zstart, s->p.zend, dz, nz, k - 1);
                    double acc = 0.0;
                    for (int ii = 1; ii <= ug; ++ii)
                        for (int jj = 1; jj <= ug; ++jj)
                            for (int kk = 1; kk <= ug; ++kk) {
                                double x = 0.5 * (xc + xm) + (double)ii * dx / (double)ug;
                                double y = 0.5 * (yc + ym) + (double)jj * dy / (double)ug;
                                double z = 0.5 * (zc + zm) + (double)kk * dz / (double)ug;
                                double inc = 1.0 / ((double)ug * (double)ug * (double)ug);
                                if (shape == 2) {
                                    if (sqrt(sq(x - c035) + sq(y - c035) + sq(z - c035)) <= c015) acc = acc + inc;
                                } else {
                                    if (sqrt(sq(x - 0.5) + sq(y - 0.75)) <= c015) acc = acc + inc;
                                    if (shape == 0 && fabs(x - 0.5) <= c0025 && y < c08 &&
                                        sq(x - 0.5) + sq(y - 0.75) <= r2)
                                        acc = acc - inc;
                                }
                            }
                    AT(vof, i, j, k) = acc;
                }
                if (shape == 0) {
                    AT(phi, i, j, k) = AT(vof, i, j, k) * 2.0 - 1.0;   /* Disk problem_init.f90:74 */
                } else if (shape == 1) {
                    AT(phi, i, j, k) = -sqrt(sq(xc - 0.5) + sq(yc - 0.75)) + c015;
                } else {
                    AT(phi, i, j, k) = -sqrt(sq(xc - c035) + sq(yc - c035) + sq(zc - c035)) + c015;
                }
            }
    orc_bc(s, phi);
    orc_bc(s, vof);
}

void orc_store_initial(orc_state *s)
{
    memcpy(s->f[ORC_PHI0], s->f[ORC_PHI], sizeof(double) * s->nbox);
    orc_ls_mv(s);
    s->ivol = s->vol;
    s->imass = s->mass;
}

void orc_mpls_step(orc_state *s, int field, double t2s, int t2red)
{
    /* main.f90:19-21: final step lands on t2s */
    if (s->time + s->p.dt > t2s) s->p.dt = t2s - s->time + 1.0e-10;
    s->time = s->time + s->p.dt;
    s->iter = s->iter + 1;
    if (field >= 0) orc_velocity(s, field, s->time, t2s);
    orc_rk3_advect(s);
    s->last_redis_iters = 0;
    if (s->iter % t2red == 0) orc_redis(s, 1, -1.0);
    orc_mass_fix(s);
    orc_ls_mv(s);
}

int orc_combine_markers(orc_state *s)
{
    /* interface_solver.f90:85-111, over the whole box including ghosts */
    const double *m1 = s->f[ORC_MK1], *m2 = s->f[ORC_MK2];
    double *phi = s->f[ORC_PHI];
    int reconstruct = 0;
    for (long q = 0; q < s->nbox; ++q) {
        const double tmp1 = m1[q], tmp2 = m2[q];
        reconstruct = reconstruct || (tmp1 * tmp2 > 0.0 && tmp1 + tmp2 > 0.0);   /* :96 */
        if (tmp1 * tmp2 > 0.0) {
            if (tmp1 + tmp2 > 0.0) phi[q] = fmin(tmp1, tmp2);                    /* :99 */
            else phi[q] = fmax(tmp1, tmp2);                                       /* :101 */
        } else {
            phi[q] = fmax(tmp1, tmp2);                                            /* :104 */
        }
    }
    orc_bc(s, phi);   /* :111 */
    return reconstruct;
}

void orc_mpls_step_markers(orc_state *s, int field, double t2s, int t2red)
{
    /* main.f90:17-26 */
    if (s->time + s->p.dt > t2s) s->p.dt = t2s - s->time + 1.0e-10;
    s->time = s->time + s->p.dt;
    s->iter = s->iter + 1;
    if (field >= 0) orc_velocity(s, field, s->time, t2s);
    s->last_redis_iters = 0;
    if (!s->merged) {                                                       /* interface_solver.f90:34 */
        for (int mid = 1; mid <= 2; ++mid) {
            /* :38-42: the marker's initial mass stands in for the global one */
            const double keep = s->imass;
            s->imass = mid == 1 ? s->imass1 : s->imass2;
            double *mk = s->f[mid == 1 ? ORC_MK1 : ORC_MK2];
            memcpy(s->f[ORC_PHI], mk, (size_t)s->nbox * sizeof(double));   /* :44-56 */
            orc_rk3_advect(s);                                              /* :60 */
            if (s->iter % t2red == 0) orc_redis(s, 1, -1.0);                /* :61 */
            orc_mass_fix(s);                                                /* :62 */
            memcpy(mk, s->f[ORC_PHI], (size_t)s->nbox * sizeof(double));   /* :64-76 */
            s->imass = keep;                                                /* :80-81 */
        }
        if (orc_combine_markers(s)) {                                       /* :85-116 */
            orc_redis(s, 1, -1.0);                                          /* :119 */
            s->merged = 1;                                                  /* :120 */
        }
    } else {
        orc_rk3_advect(s);                                                  /* :125-127 */
        if (s->iter % t2red == 0) orc_redis(s, 1, -1.0);
        orc_mass_fix(s);
    }
    orc_ls_mv(s);                                                           /* main.f90:34 */
}

/* ------------------------------------------------------------------------
 * CLSVOF
 * ---------------------------------------------------------------------- */
void orc_velocity_faces(orc_state *s, int field, double time, double t2s)
{
    const double pi = acos(-1.0);
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz, ks = s->ks, ke = s->ke;
    double *uf = s->f[ORC_UF], *vf = s->f[ORC_VF], *wf = s->f[ORC_WF];
    const double ct = cos(pi * time / t2s), omega = 2.0 * pi / t2s;
    for (int k = ks; k <= ke; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double x = coord(s->p.xstart, s->p.xend, s->p.dx, nx, i);
                double y = coord(s->p.ystart, s->p.yend, s->p.dy, ny, j);
                double xx = 0.5 * (x + coord(s->p.xstart, s->p.xend, s->p.dx, nx, i + 1));
                double yy = 0.5 * (y + coord(s->p.ystart, s->p.yend, s->p.dy, ny, j + 1));
                if (field == 0) {
                    /* Disk/src/setup/velocity.f90:88-89: the cell-centre formulas */
                    AT(uf, i, j, k) = -omega * (y - 0.5);
                    AT(vf, i, j, k) = omega * (x - 0.5);
                } else if (field == 1) {
                    /* Vortex2D/src/setup/velocity.f90:38-39 */
                    AT(uf, i, j, k) = -sq(sin(pi * xx)) * sin(2.0 * pi * y) * ct;
                    AT(vf, i, j, k) = sq(sin(pi * yy)) * sin(2.0 * pi * x) * ct;
                } else {
                    /* Vortex/src/setup/velocity.f90:31-33 */
                    double z = coord(s->p.zstart, s->p.zend, s->p.dz, nz, k);
                    double zz = 0.5 * (z + coord(s->p.zstart, s->p.zend, s->p.dz, nz, k + 1));
                    AT(uf, i, j, k) = 2.0 * sq(sin(pi * xx)) * sin(2.0 * pi * y) * sin(2.0 * pi * z) * ct;
                    AT(vf, i, j, k) = -sin(2.0 * pi * x) * sq(sin(pi * yy)) * sin(2.0 * pi * z) * ct;
                    AT(wf, i, j, k) = -sin(2.0 * pi * x) * sin(2.0 * pi * y) * sq(sin(pi * zz)) * ct;
                }
            }
    orc_bc(s, uf); orc_bc(s, vf); orc_bc(s, wf);
}

/* WLIC-THINC flux through the face between cells c and c + e along d
 * (clsvof.f90:26-63 for x, :91-128 for y, :191-228 for z): the cell upwind
 * of the face velocity uf gives vof; a full or empty cell is carried by the
 * face velocity, otherwise the THINC profile's flux (slope alpha from the
 * cells either side of it, ib = max(1, index - 1)) is weighted with the
 * upwind one by |n_d| / (|n_x| + |n_y| + |n_z|). */
static double wlic_flux(orc_state *s, int d, int i, int j, int k)
{
    const double eps = 1.0e-10, beta = 2.3, dt = s->p.dt;
    const double *vof = s->f[ORC_VOF];
    const double *vel = s->f[d == 0 ? ORC_UF : d == 1 ? ORC_VF : ORC_WF];
    const double *N[3] = {s->f[ORC_NX], s->f[ORC_NY], s->f[ORC_NZ]};
    const double dd = d == 0 ? s->p.dx : d == 1 ? s->p.dy : s->p.dz;
    const double uf = AT(vel, i, j, k);
    int c[3] = {i, j, k};
    double isgn;
    if (uf > 0.0) {
        isgn = 1.0;
    } else {
        c[d] += 1;
        isgn = 0.0;
    }
    const double v = AT(vof, c[0], c[1], c[2]);
    if (v > 1.0 - eps || v < eps) return v * uf * dt;
    int b[3] = {c[0], c[1], c[2]}, a[3] = {c[0], c[1], c[2]};
    b[d] = c[d] - 1 > 1 ? c[d] - 1 : 1;   /* ib = max(1, ii - 1) */
    a[d] = c[d] + 1;
    const double alpha = AT(vof, b[0], b[1], b[2]) < AT(vof, a[0], a[1], a[2]) ? 1.0 : -1.0;
    const double a1 = exp(beta * (2.0 * v - 1.0) / alpha);
    const double a3 = exp(beta);
    const double xc = 0.5 / beta * log((a3 * a3 - a1 * a3) / (a1 * a3 - 1.0));
    const double a4 = cosh(beta * (isgn - uf * dt / dd - xc));
    const double a5 = cosh(beta * (isgn - xc));
    double w = fabs(AT(N[0], c[0], c[1], c[2])) + fabs(AT(N[1], c[0], c[1], c[2])) + fabs(AT(N[2], c[0], c[1], c[2]));
    w = fabs(AT(N[d], c[0], c[1], c[2])) / w;
    double f = 0.5 * (uf * dt - alpha * dd / beta * log(a4 / a5));
    f = f * w + (1.0 - w) * v * uf * dt;
    return f;
}

void orc_vof_wlic(orc_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, ks = s->ks, ke = s->ke;
    const double dt = s->p.dt;
    double *vof = s->f[ORC_VOF];
    const double *old = s->f[ORC_VOF_OLD];
    orc_surface_norms(s);   /* clsvof.f90:13 */
    const int ndir = s->p.dim == 3 ? 3 : 2;
    for (int d = 0; d < ndir; ++d) {
        double *F = s->f[d == 0 ? ORC_SX1 : d == 1 ? ORC_SY1 : ORC_SZ1];   /* tdata%x/y/z%s1 */
        const double *vel = s->f[d == 0 ? ORC_UF : d == 1 ? ORC_VF : ORC_WF];
        const double dd = d == 0 ? s->p.dx : d == 1 ? s->p.dy : s->p.dz;
        /* faces: index 0 .. n along d, interior along the others */
        for (int k = (d == 2 ? ks - 1 : ks); k <= ke; ++k)
            for (int j = (d == 1 ? 0 : 1); j <= ny; ++j)
                for (int i = (d == 0 ? 0 : 1); i <= nx; ++i) AT(F, i, j, k) = wlic_flux(s, d, i, j, k);
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const int im = d == 0 ? i - 1 : i, jm = d == 1 ? j - 1 : j, km = d == 2 ? k - 1 : k;
                    const double sflux = (AT(F, i, j, k) - AT(F, im, jm, km)) / dd;
                    AT(vof, i, j, k) = AT(vof, i, j, k) - sflux +
                                       AT(old, i, j, k) * dt * (AT(vel, i, j, k) - AT(vel, im, jm, km)) / dd;
                    if (d == 2) AT(vof, i, j, k) = fmin(1.0, fmax(0.0, AT(vof, i, j, k)));   /* :256 */
                }
        orc_bc(s, vof);
    }
    if (ndir == 2) {
        /* Vortex2D clsvof.f90:241-262: the z loop keeps only the clamp */
        for (int k = ks; k <= ke; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) AT(vof, i, j, k) = fmin(1.0, fmax(0.0, AT(vof, i, j, k)));
        orc_bc(s, vof);
    }
}

void orc_clsvof_recon(orc_state *s)
{
    double *phi = s->f[ORC_PHI], *vof = s->f[ORC_VOF], *tmp = s->f[ORC_VOF_TMP];
    for (long q = 0; q < s->nbox; ++q) {   /* :283-292, the whole box */
        tmp[q] = phi[q];
        phi[q] = 2.0 * vof[q] - 1.0;
    }
    orc_redis(s, 0, 5.0 * s->p.ls_wid);   /* :297 */
    for (long q = 0; q < s->nbox; ++q)    /* :305-311 */
        if (fabs(tmp[q]) > 1.5 * s->p.ls_wid) phi[q] = tmp[q];
    orc_bc(s, phi);
    orc_redis(s, 1, -1.0);                 /* :317 */
}

void orc_clsvof_step(orc_state *s, int field, double t2s, int t2red)
{
    if (s->time + s->p.dt > t2s) s->p.dt = t2s - s->time + 1.0e-10;   /* main.f90:19-21 */
    s->time = s->time + s->p.dt;
    s->iter = s->iter + 1;
    if (field >= 0) {   /* main.f90:26 */
        orc_velocity(s, field, s->time, t2s);
        orc_velocity_faces(s, field, s->time, t2s);
    }
    memcpy(s->f[ORC_VOF_OLD], s->f[ORC_VOF], sizeof(double) * s->nbox);   /* p%switch, main.f90:27 */
    orc_vof_wlic(s);                        /* interface_solver.f90:236 */
    orc_rk3_advect(s);                      /* :238 */
    s->last_redis_iters = 0;
    if (s->iter % t2red == 0) orc_clsvof_recon(s);   /* :240-242 */
    orc_ls_mv(s);                           /* main.f90:33 */
}
