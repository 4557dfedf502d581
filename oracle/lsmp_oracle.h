/*
 * lsmp_oracle.h -- CPU restatement (FP64, scalar C) of the reference's
 * mass-preserving level-set transport step and of its CLSVOF step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file and lsmp_oracle.c are the checker for
 * the HIP library: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load them.  The product never links or calls them.
 *
 * Layout follows the reference storage contract exactly: every field is a
 * column-major (i fastest) FP64 box over (1-ghc : n+ghc) in each direction
 * with ghc = 3 (Disk/lib/data/branches.f90:134-168).  In 2D the z extent is
 * 1-3 : 1+3 (7 planes) because ks = ke = 1 (Disk/lib/data/branches.f90:126-127).
 *
 * Every routine cites the reference file:line it restates; paths are
 * relative to the reference checkout (Vortex/ for 3D forms, Disk/ for 2D).
 */
#ifndef LSMP_ORACLE_H
#define LSMP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int nx, ny, nz;      /* interior cells; nz = 1 when dim == 2 */
    int dim;             /* 2 or 3 */
    double dx, dy, dz;   /* dz = 1 in 2D (Disk/lib/data/tree.f90:252) */
    double dt, rdt;      /* already multiplied by dx (tree.f90:278-279) */
    double ls_wid;       /* already multiplied by dx (tree.f90:280) */
    double rho_12;
    int inverse;
    double xstart, xend, ystart, yend, zstart, zend;
} orc_params;

typedef struct orc_state orc_state;

/* field selectors for orc_field() */
enum {
    ORC_PHI = 0, ORC_PHI_TMP, ORC_PHI0, ORC_U, ORC_V, ORC_W,
    ORC_HEAVY, ORC_DELTA, ORC_SIGN, ORC_SIGN_TMP, ORC_GRAD,
    ORC_SX1, ORC_SX2, ORC_SY1, ORC_SY2, ORC_SZ1, ORC_SZ2,
    ORC_SSX1, ORC_SSX2, ORC_SSY1, ORC_SSY2, ORC_SSZ1, ORC_SSZ2,
    ORC_S3, ORC_L1, ORC_L2, ORC_L3,
    ORC_NX, ORC_NY, ORC_NZ, ORC_NXX, ORC_NYY, ORC_NZZ,
    ORC_VOF,
    ORC_MK1, ORC_MK2,    /* the two markers' level-set fields (marker(mid)%lsf) */
    ORC_VOF_OLD,         /* vof%old: vof at the step's start (p%switch) */
    ORC_UF, ORC_VF, ORC_WF,   /* face velocities vel%x/y/z%old (x face i + 1/2 at index i) */
    ORC_VOF_TMP,         /* vof%tmp: clsvof_recon keeps the old phi there */
    ORC_NFIELDS
};

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *s);
double *orc_field(orc_state *s, int which);   /* box array, (nx+6)*(ny+6)*nzb */
long orc_box_size(const orc_state *s);
void orc_box_dims(const orc_state *s, int *nxb, int *nyb, int *nzb);

/* scalars */
double orc_get(const orc_state *s, const char *name);
void orc_set(orc_state *s, const char *name, double v);

/* ---- line-level routines (arrays are 0-based C arrays) ---- */
/* TDMA.f90:1-36; arrays of length n, modified in place like the reference */
void orc_solve_tridiagonal(double *A, double *B, double *C, double *S, double *X, int n);
/* level_set_rk3_solver.f90:288-345 / 347-414.  Line arrays hold the box line
 * (n interior + 3 ghosts each side, so length n+6); interior index i=1..n is
 * stored at [i+2]. */
void orc_wenojs_flux_split(const double *f, const double *g, double *fp, double *gm, int n);
void orc_crweno_flux_split(const double *f, const double *g, double *fp, double *gm, int n);
/* level_set_rk3_redis.f90:456-471 */
double orc_phyn(double a, double b, double c, double d);
/* ccd_solvers.f90:80-163, 291-310, 412-427; matrix_solver.f90:6-78.
 * f, fx, fxx have length n (the full box line). */
void orc_ccd_line(int n, double dx, const double *f, double *fx, double *fxx);
/* upwind CCD first derivative, ccd_solvers.f90:165-214, 312-345, 428-438 */
void orc_uccd_line(int n, double dx, const double *f, const double *u, double *fx);

/* ---- field-level routines, on the state ---- */
void orc_bc(orc_state *s, double *dat);                   /* branches.f90:275-320 */
void orc_ls_funs(orc_state *s);                           /* ns_solver/others.f90:1-45 */
void orc_ls_mv(orc_state *s);                             /* ns_solver/others.f90:172-213 */
void orc_rk3_source(orc_state *s);                        /* level_set_rk3_solver.f90:104-221 */
void orc_rk3_advect(orc_state *s);                        /* level_set_rk3_solver.f90:1-94 */
void orc_rk3_source_uccd(orc_state *s);                   /* level_set_rk3_solver.f90:223-286 */
/* orc_set(s, "scheme", 1) makes orc_rk3_advect use the UCCD RHS (the
 * reference's level_set_rk3_source selector, :96-102). */
void orc_redis_gradient(orc_state *s);                    /* level_set_rk3_redis.f90:266-369 */
void orc_redis_lambda(orc_state *s, int btn);             /* level_set_rk3_redis.f90:371-454 */
/* level_set_rk3_redis.f90:1-79; timestop < 0 selects the reference default.
 * returns the number of pseudo-steps run. */
int orc_redis(orc_state *s, int btn, double timestop);
void orc_surface_norms(orc_state *s);                     /* derivatives.f90:1-50, 96-116 */
int orc_mass_fix(orc_state *s);                           /* mass_preserving_level_set.f90:1-74 */
double orc_interface_l1(orc_state *s);                    /* output.f90:169-204, 233-248 */
/* output.f90:250-397 with rho_mu (ns_solver/others.f90:47-80): out[11] = A, v, px/v, py/v,
   pz/v, xc/v, yc/v, zc/v, Ke, Po, Q (unscaled); vof_marker: method 3's marker */
void orc_energy_momentum(orc_state *s, double mu_12, int vof_marker, double *out);

/* ---- inputs (what the reference's setup routines produce) ---- */
/* src/setup/velocity.f90:1-60 (3D Vortex form and 2D Vortex2D form),
 * Disk/src/setup/velocity.f90:63-114 (rigid rotation).  field: 0 rotation,
 * 1 vortex2d, 2 leveque3d.  Writes u,v,w (nvel) boxes incl. ghosts. */
void orc_velocity(orc_state *s, int field, double time, double t2s);
/* src/setup/problem_init.f90:32-79.  shape: 0 zalesak disk (phi = 2 vof - 1),
 * 1 circle r=0.15 at (0.5,0.75), 2 sphere r=0.15 at (0.35,0.35,0.35).
 * with_vof != 0 also runs the 30^3 sub-sampling VOF. */
void orc_initial_phi(orc_state *s, int shape, int with_vof);
/* problem_init.f90:104-150: phi0 <- phi (tmp2), then imass/ivol via ls_mv */
void orc_store_initial(orc_state *s);
/* one full MPLS time step (main.f90:17-35 + interface_solver.f90:123-129):
 * dt clamp, t += dt, iter += 1, velocity at t, rk3, redis if
 * mod(iter,t2red)==0, mass fix, ls_mv.  field as in orc_velocity; field < 0
 * keeps the current u,v,w. */
void orc_mpls_step(orc_state *s, int field, double t2s, int t2red);
/* interface_solver.f90:85-111: phi <- merge of MK1 and MK2 over the whole
 * box (same sign: the smaller inside both, the larger outside both; else the
 * larger); returns 1 when some cell is inside both markers. */
int orc_combine_markers(orc_state *s);
/* one MPLS step of the two-marker branch (main.f90:17-34 with
 * interface_solver.f90:34-128): while "merged" is 0, each marker in turn is
 * copied into phi, advected, reinitialised (iter mod t2red == 0) and mass
 * corrected against its own initial mass ("imass1", "imass2"), and copied
 * back; phi then becomes the merge of the two, reinitialised once and
 * "merged" set when they overlap.  Once merged, the one-field step.  ls_mv at
 * the end. */
void orc_mpls_step_markers(orc_state *s, int field, double t2s, int t2red);

/* ---- CLSVOF (coupled level set and VOF, the Vortex deck's default) ---- */
/* face velocities vel%x/y/z (src/setup/velocity.f90 of each case): the
 * LeVeque and vortex fields at the face midpoints xx = (x(i) + x(i+1)) / 2
 * (Vortex/src/setup/velocity.f90:23-34; Vortex2D/src/setup/velocity.f90:
 * 23-24, 38-39); the rotation's faces take the cell-centre formulas
 * (Disk/src/setup/velocity.f90:89-93).  Writes UF, VF, WF incl. ghosts. */
void orc_velocity_faces(orc_state *s, int field, double time, double t2s);
/* vof_wlic_solver (Vortex/src/interface_solver/clsvof.f90:1-273): CCD
 * normals, then per direction the WLIC-THINC face fluxes (beta = 2.3) and
 * the update vof -= dF/dd - vof_old dt du/dd; the last direction clamps to
 * [0, 1].  In 2D the z fluxes are dropped and the clamp stays
 * (Vortex2D/src/interface_solver/clsvof.f90:183-256). */
void orc_vof_wlic(orc_state *s);
/* clsvof_recon (clsvof.f90:275-324): phi <- 2 vof - 1, reinitialised with
 * btn = 0 to T = 5 ls_wid, the old phi kept where |phi_old| > 1.5 ls_wid,
 * then the btn = 1 reinitialisation. */
void orc_clsvof_recon(orc_state *s);
/* one CLSVOF step of the merged branch (main.f90:17-34 with
 * interface_solver.f90:234-244): dt clamp, t += dt, iter += 1, node and face
 * velocity at t (field < 0 keeps them), vof_old <- vof, vof_wlic, rk3
 * advection of phi, clsvof_recon if mod(iter, t2red) == 0, ls_mv. */
void orc_clsvof_step(orc_state *s, int field, double t2s, int t2red);

#ifdef __cplusplus
}
#endif
#endif
