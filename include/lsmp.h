/*
 * lsmp.h -- C ABI of the MI355X (gfx950) level-set transport library.
 *
 * The library replaces the reference's hot path: TVD-RK3 CRWENO5 advection,
 * PDE reinitialisation and the global mass-preserving correction of the MPLS
 * method.  In the reference these are argument-less internal Fortran
 * subroutines acting on the globals of module `all` (Vortex/src/main.f90:1-8);
 * the entry points below are what a Fortran driver binds through
 * ISO_C_BINDING in their place (one entry point per reference call site).
 *
 * Conventions
 *  - Host arrays are FP64 boxes in the reference storage contract: column-major
 *    (i fastest), extents (1-ghc : n+ghc) per direction with ghc = 3; in 2D the
 *    z extent is 7 planes (Disk/lib/data/branches.f90:134-136).
 *  - The library owns every device buffer; phi stays resident in HBM between
 *    calls.  Host pointers are never retained past a call.
 *  - Every function returns 0 on success or a negative LSMP_E* code; the text
 *    is available from lsmp_error_string().  The reference stops the program
 *    on the corresponding conditions (Disk/lib/data/tree.f90:304-306,
 *    branches.f90:129-132, 314-315).
 *  - One host thread per handle; calls on a handle are not re-entrant, which is
 *    the reference's own call pattern.
 */
#ifndef LSMP_H
#define LSMP_H

#ifdef __cplusplus
extern "C" {
#endif

#define LSMP_OK 0
#define LSMP_E_INVALID (-1)    /* bad parameter / size */
#define LSMP_E_HIP (-2)        /* HIP runtime error */
#define LSMP_E_NAN (-3)        /* non-finite value detected */
#define LSMP_E_EMPTY_BAND (-4) /* mass-fix denominator is zero: no cell in the band */
#define LSMP_E_STATE (-5)      /* call out of order (e.g. no velocity set) */

/* arithmetic of the bulk arrays */
#define LSMP_FP64 0            /* plain form, drop-in parity with the reference */
#define LSMP_FP32 1            /* FP32 bulk; mass sums, lambda and area in FP64 */

/* prescribed velocity fields for lsmp_set_velocity_analytic */
#define LSMP_VEL_ROTATION 0    /* Disk/src/setup/velocity.f90:63-114 */
#define LSMP_VEL_VORTEX2D 1    /* Vortex2D/src/setup/velocity.f90:35-36 */
#define LSMP_VEL_LEVEQUE3D 2   /* Vortex/src/setup/velocity.f90:27-29 */

typedef struct lsmp_params {
    int nx, ny, nz;          /* interior cells (glb%node_x..z); nz = 1 in 2D */
    int ghc;                 /* ghost width of host boxes; must be 3 */
    int dim;                 /* 2 or 3 */
    double dx, dy, dz;       /* dz = 1 in 2D (Disk/lib/data/tree.f90:252) */
    double dt;               /* physical step, already scaled by dx (tree.f90:278) */
    double rdt;              /* pseudo-time step, scaled by dx (tree.f90:279) */
    double ls_wid;           /* interface half width, scaled by dx (tree.f90:280) */
    double rho_12;           /* density ratio (deck line 50) */
    int inverse;             /* inverse marker function (deck line 52) */
    int precision;           /* LSMP_FP64 | LSMP_FP32 */
    double xstart, xend, ystart, yend, zstart, zend;   /* domain (deck 18-22) */
} lsmp_params;

typedef struct lsmp_stats {
    int last_redis_iters;    /* pseudo-steps of the last lsmp_reinit */
    int last_mass_iters;     /* iterations of the last lsmp_mass_fix */
    double last_red_error;   /* max |dphi| of the last pseudo-step */
    double t_advect_ms;      /* accumulated device time per phase (HIP events) */
    double t_reinit_ms;
    double t_massfix_ms;
    long long n_advect, n_reinit, n_massfix;  /* calls */
} lsmp_stats;

typedef struct lsmp_ctx *lsmp_handle;

/* Set-up: replaces job_init's allocation and the CCD factorisation done once
 * per solver (Vortex/lib/data/branches.f90:97-174, ccd_solvers.f90:382-410). */
int lsmp_init(const lsmp_params *params, lsmp_handle *out);
void lsmp_finalize(lsmp_handle h);
const char *lsmp_error_string(int code);

/* phi in / out as a host box (ghosts are filled on download by the
 * zero-gradient rule, branches.f90:275-320). */
int lsmp_upload_phi(lsmp_handle h, const double *phi_box);
int lsmp_download_phi(lsmp_handle h, double *phi_box);
/* Keep the current phi as phi0 = phi%tmp2, the target of the L1 shape metric
 * (problem_init.f90:104-113 -> roots.f90:150-166). */
int lsmp_store_phi0(lsmp_handle h);

/* Node velocity nvel%old used by the advection (level_set_rk3_solver.f90:118-125).
 * w is ignored in 2D. */
int lsmp_set_velocity(lsmp_handle h, const double *u_box, const double *v_box, const double *w_box);
/* The same velocity evaluated on the device from the reference's formulas at
 * time t (velocity.f90; main.f90:23-28 evaluates it at t^{n+1}). */
int lsmp_set_velocity_analytic(lsmp_handle h, int field, double t, double t2s);

/* One TVD-RK3 advection step with step dt (level_set_rk3_solver.f90:1-94),
 * with the RHS of the selected scheme: CRWENO5 flux split (default,
 * :104-221) or upwind CCD (:223-286). */
int lsmp_rk3_advect(lsmp_handle h, double dt);
/* The reference's level_set_rk3_source selector (:96-102), a source edit
 * there: LSMP_SCHEME_CRWENO (source_weno) or LSMP_SCHEME_UCCD (source_ccd,
 * -u . grad phi with the "uccd" solver of ccd_solvers.f90:165-214, 312-345,
 * 428-438). */
#define LSMP_SCHEME_CRWENO 0
#define LSMP_SCHEME_UCCD 1
int lsmp_set_advection_scheme(lsmp_handle h, int scheme);

/* Reinitialisation toward |grad phi| = 1 (level_set_rk3_redis.f90:1-79).
 * btn = 1: sign from the smoothed Heaviside, constrained lambda; btn = 0:
 * raw-phi sign, phi /= max|grad phi|, no lambda (Disk initialisation).
 * timestop < 0 selects the reference default (2 ls_wid, or 1.5 x domain
 * extent for btn = 0); otherwise it is the optional argument T. */
int lsmp_reinit(lsmp_handle h, int btn, double timestop, double *red_error, int *n_pseudo);

/* LS mass and volume sums of the current phi (ns_solver/others.f90:172-213). */
int lsmp_mass_vol(lsmp_handle h, double *mass, double *vol);

/* The sums of lsmp_mass_vol on a caller's device field, allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory or synchronise (code under HIP graph capture, tight serving loops).
 *   phi     device FP64 interior cells, (k*ny + j)*nx + i (nz = 1 in 2D)
 *   dx..dz  cell sizes (2D: dz = 1, tree.f90:252); ls_wid, rho_12, inverse
 *           as in lsmp_params
 *   out     device double[2] <- (mass, vol), stream-ordered (no host wait)
 *   stream  hipStream_t, NULL for the default stream
 * Results equal lsmp_mass_vol's on the same field bit for bit.  Scratch: a
 * static device array of LSMP_NOALLOC_PARTIALS partial triples (24 KiB) per
 * GPU, held until the process ends; its slot count caps the reduction's
 * block count, so every grid fits.  Calls that overlap on one GPU share that
 * array and race. */
#define LSMP_NOALLOC_PARTIALS 1024
int lsmp_mass_vol_noalloc(const double *phi, int nx, int ny, int nz, int dim, double dx, double dy, double dz,
                          double ls_wid, double rho_12, int inverse, double *out, void *stream);

/* Global mass correction with target imass (mass_preserving_level_set.f90:1-74).
 * mass receives the mass measured at the start of the last iteration (the
 * value the exit test used), n_iter the iterations run. */
int lsmp_mass_fix(lsmp_handle h, double imass, double *mass, int *n_iter);

/* Band lines in the mass correction (no counterpart in the reference, whose
 * surface_norms solves every line, derivatives.f90:11-50): the CCD normals
 * solved only on the lines that cross the interface band |phi| < ls_wid, the
 * only cells where the correction delta |grad phi| and its sum are nonzero.
 * mode 0: every line; 1 (the default): band lines on 3D grids with more than
 * 4096 lines in a direction and lines of at most 540 cells, where they pay,
 * every line elsewhere; 2: band lines always.  The modes agree to the drop-in
 * tolerance (the listed lines take the 32-lanes-per-line solve). */
int lsmp_set_band_lines(lsmp_handle h, int mode);
/* Lines listed by the last mass correction along x, y and z (z: 0 in 2D);
 * all -1 when it solved every line. */
int lsmp_band_line_counts(lsmp_handle h, int counts[3]);

/* Narrow band for the chosen forms (no counterpart in the reference, which
 * transports and reinitialises every cell).  width_cells = W / dx > 0 (3D
 * only, at least 4): lsmp_rk3_advect (CRWENO) first clamps phi to [-W, W],
 * then solves only the lines that cross the tube |phi| < W and updates only
 * the cells whose three lines do; lsmp_reinit updates only the workgroups of
 * 32 x 8 x 8 cells that hold a cell |phi| < W; every other cell keeps its
 * value.  Results differ from the full-domain method away from the interface
 * (the far field is the constant +-W), so the FP64 drop-in leaves it off
 * (width_cells = 0, the default).  With LSMP_SCHEME_UCCD the advection runs on
 * every cell and only the reinitialisation is narrowed. */
int lsmp_set_narrow_band(lsmp_handle h, double width_cells);

/* L1 shape error sum |H(phi0/eps) - H(phi/eps)| / (nx ny nz)
 * (output.f90:169-204, 226, 233-248). */
int lsmp_interface_l1(lsmp_handle h, double *l1);

/* The per-step energy and momentum integrals of energy_and_momentum
 * (output.f90:250-397, called every step by output(), :33) with rho_mu
 * (ns_solver/others.f90:47-80): the nine first CCD derivatives of the
 * velocity, then out[11] = A (the surface-area estimate, already / dx), v
 * (sum marker dV), px, py, pz (sum marker u dV / v ...), xc, yc, zc (the
 * centroid), Ke, Po, Q (kinetic and potential energy and the viscous
 * dissipation rate over the cells with marker > 0.99) -- before the deck's
 * Fr, We, Re and energy_unit scaling (:374-386), which the caller applies.
 * marker = H(phi), or the VOF with vof_marker != 0 (method 3); mu_12 is the
 * deck's viscosity ratio.  Needs the velocity (and the VOF); the first call
 * allocates 9 fields for the derivatives. */
int lsmp_energy_momentum(lsmp_handle h, double mu_12, int vof_marker, double *out);

/* Device-time accounting per phase (the reference's ls_adv / ls_red timers,
 * level_set_rk3_solver.f90:91-92, level_set_rk3_redis.f90:76-77). */
int lsmp_get_stats(lsmp_handle h, lsmp_stats *out);
/* Per-kernel timing with HIP events on the library's stream (off by default).
 * lsmp_kernel_time reports the launches and summed milliseconds of one kernel
 * by name ("adv_x", "adv_y", "adv_z", "redis_stage", "ccd_x", ...). */
int lsmp_set_profiling(lsmp_handle h, int on);
int lsmp_kernel_time(lsmp_handle h, const char *name, long long *launches, double *ms);
/* Block until all work queued by the handle has finished. */
int lsmp_synchronize(lsmp_handle h);

/* Per-interface mass (no counterpart in the reference, which conserves and
 * reports the total only): the 6-connected components of the interior cells
 * with phi > phi_core, numbered by their smallest cell index; each cell with
 * phi_min < phi <= phi_core joins the component that reaches it in the fewest
 * face steps through such cells (the lower number on ties; a cell no
 * component reaches joins none).  phi_core = 0, phi_min = -ls_wid keeps two
 * interfaces apart as long as their zero level sets do not touch;
 * phi_core = phi_min labels the region phi > phi_min as such.  Each
 * component gets the FP64 level-set mass sum rho(H) H dV of its cells (the
 * integrand of ls_mv, ns_solver/others.f90:172-213).  query_cells (interior
 * flat indices (k ny + j) nx + i) receive their component number or -1;
 * comp_mass gets the first maxcomp masses, *ncomp the number of components.
 * LSMP_E_INVALID if phi_min > phi_core. */
int lsmp_component_masses(lsmp_handle h, double phi_core, double phi_min, int nquery, const long long *query_cells,
                          int *query_comp, int maxcomp, double *comp_mass, int *ncomp);

/* ------------------------------------------------------------------------
 * Two-marker MPLS branch (mpls, interface_solver.f90:34-121): while the case
 * is not merged, each marker (marker(mid)%lsf) is transported on its own --
 * lsmp_marker_load(mid), lsmp_rk3_advect, lsmp_reinit(1), lsmp_mass_fix
 * against that marker's initial mass, lsmp_marker_store(mid) -- then
 * lsmp_combine_markers forms phi from the two and reports whether some cell is
 * inside both, in which case the caller reinitialises phi once and continues
 * with phi alone (glb%merged).
 * ------------------------------------------------------------------------ */
/* upload / download marker mid (1 or 2), a box like lsmp_upload_phi's
 * (problem_init.f90:83-84 set them); LSMP_E_STATE for a marker never set */
int lsmp_set_marker(lsmp_handle h, int mid, const double *box);
int lsmp_get_marker(lsmp_handle h, int mid, double *box);
/* phi <- marker mid (interface_solver.f90:44-56); marker mid <- phi (:64-76) */
int lsmp_marker_load(lsmp_handle h, int mid);
int lsmp_marker_store(lsmp_handle h, int mid);
/* phi <- merge of the two markers (:85-111): where both have the same sign,
 * the smaller inside both and the larger outside both, elsewhere the larger;
 * *reconstruct = 1 when some cell is inside both (:96).  LSMP_E_STATE unless
 * both markers are set. */
int lsmp_combine_markers(lsmp_handle h, int *reconstruct);

/* ------------------------------------------------------------------------
 * CLSVOF (coupled level set and volume of fluid; clsvof, interface_solver.f90:
 * 133-244, merged branch :234-244).  A step as the reference runs it, after
 * the driver set the velocities at t^{n+1} (main.f90:23-27):
 * lsmp_set_velocity[_analytic] and lsmp_set_face_velocity[_analytic],
 * lsmp_store_vof_old, lsmp_vof_wlic(dt), lsmp_rk3_advect(dt), and every t2red
 * steps lsmp_clsvof_recon.  The CLSVOF fields are allocated on the first of
 * these calls.  Not available with the narrow band on (LSMP_E_STATE).
 * ------------------------------------------------------------------------ */
/* vof%now as a box like lsmp_upload_phi's (problem_init.f90 fills it);
 * download: LSMP_E_STATE before an upload */
int lsmp_upload_vof(lsmp_handle h, const double *vof_box);
int lsmp_download_vof(lsmp_handle h, double *vof_box);
/* vof%old <- vof%now, the VOF part of p%switch (Vortex/lib/data/tree.f90:383-392) */
int lsmp_store_vof_old(lsmp_handle h);
/* vof%tmp2 <- vof%now, the initial VOF of the shape error (problem_init's loc%init) */
int lsmp_store_vof0(lsmp_handle h);
/* VOF half of ls_mv (ns_solver/others.f90:215-247): volume sum vof dv and
 * mass sum rho(vof) vof dv (p%glb%volv, p%glb%massv) */
int lsmp_vof_mass_vol(lsmp_handle h, double *mass, double *vol);
/* VOF part of calculate_interface_loss (output.f90:178-204, "VOF Int"):
 * sum |vof0 - vof| / (nx ny nz); LSMP_E_STATE before lsmp_store_vof0 */
int lsmp_vof_interface_l1(lsmp_handle h, double *l1);
/* Face velocities vel%x/y/z%old: index i of the x box is the face i + 1/2
 * (and likewise for y, z); the ghost faces are the zero-gradient fill of the
 * first and last face, as the reference's bc sets them.  w is ignored in 2D. */
int lsmp_set_face_velocity(lsmp_handle h, const double *u_box, const double *v_box, const double *w_box);
/* The same from the prescribed fields at time t: LeVeque and vortex at the
 * face midpoints (Vortex/src/setup/velocity.f90:20-33, Vortex2D/src/setup/
 * velocity.f90:38-39), the rotation with the cell-centre formulas
 * (Disk/src/setup/velocity.f90:89-93). */
int lsmp_set_face_velocity_analytic(lsmp_handle h, int field, double t, double t2s);
/* vof_wlic_solver (Vortex/src/interface_solver/clsvof.f90:1-273) with step
 * dt: CCD normals of phi (surface_norms), then per direction the WLIC-THINC
 * face fluxes and the update vof -= dF/dd - vof_old dt du/dd, the last
 * direction clipping to [0, 1].  LSMP_E_STATE before lsmp_upload_vof,
 * lsmp_store_vof_old or a face velocity.  Where the interface reaches a
 * boundary cell, the normal along the pass's direction at the ghost cell
 * beyond it is the boundary cell's (the reference takes the CCD derivative at
 * the ghost point). */
int lsmp_vof_wlic(lsmp_handle h, double dt);
/* clsvof_recon (clsvof.f90:275-324): phi <- 2 vof - 1, level_set_rk3_redis(0,
 * 5 ls_wid), the previous phi restored where its magnitude exceeds 1.5 ls_wid,
 * level_set_rk3_redis(1).  n_pseudo0 / n_pseudo1 (may be null) receive the
 * pseudo-steps of the two reinitialisations. */
int lsmp_clsvof_recon(lsmp_handle h, int *n_pseudo0, int *n_pseudo1);

/* ------------------------------------------------------------------------
 * Learned reinitialisation.  No counterpart in the reference: it replaces
 * most pseudo-steps of level_set_rk3_redis (level_set_rk3_redis.f90:25-74)
 * by one pass of a small network on the narrow band, followed by a few plain
 * pseudo-steps and a residual check that falls back to the plain method.
 *
 * Network: 128 inputs (the 5x5x5 stencil (phi(nb) - phi(c)) / h with the
 * x offset fastest, then phi(c) / h, then two zeros) -> 64 SiLU -> 64 SiLU
 * -> 1, output = (phi_reinit(c) - phi(c)) / h.  Parameters are FP32 masters
 * (W1[128][64], b1[64], W2[64][64], b2[64], w3[64], b3, row-major in that
 * order).  Products run on the matrix cores in split BF16: each FP32 operand
 * is hi + lo with hi = bf16(v), lo = bf16(v - hi), and a product is
 * hi*hi + lo*hi + hi*lo with FP32 accumulation (about 16 significant bits).
 * h = dx; 3-D grids need dx = dy = dz.
 * ------------------------------------------------------------------------ */
typedef struct lsmp_net *lsmp_net_handle;

/* capacity: training pairs the net can hold (0 for inference only);
 * parameters start from a seeded uniform initialisation. */
int lsmp_net_create(int capacity, unsigned seed, lsmp_net_handle *out);
void lsmp_net_destroy(lsmp_net_handle net);
int lsmp_net_param_count(void);
int lsmp_net_set_params(lsmp_net_handle net, const float *params, int n);  /* resets Adam */
int lsmp_net_get_params(lsmp_net_handle net, float *params, int n);
int lsmp_net_dataset_size(lsmp_net_handle net, int *n);
/* Run the plain reinitialisation (btn = 1, default timestop) on the solver's
 * phi and store one (pre, post) pair for every stride-th cell (ascending
 * index order) with |phi_pre| < band_cells dx until the capacity is reached.
 * phi is left reinitialised. */
int lsmp_net_collect(lsmp_net_handle net, lsmp_handle h, double band_cells, int stride, int *n_added);
/* Append n host pairs x[n][128], y[n] (FP32); drop all pairs. */
int lsmp_net_add_samples(lsmp_net_handle net, const float *x, const float *y, int n);
int lsmp_net_clear_samples(lsmp_net_handle net);
/* phi <- phi + dx * net(stencil) * clamp((|phi| - ls_wid) / ls_wid, 0, 1) on
 * |phi| < band_cells dx, nothing else (the network pass of
 * lsmp_reinit_learned on its own).  The network leaves the interface region
 * |phi| <= ls_wid to the plain sweeps and the mass correction.  *n_band
 * receives the number of band cells; n_band may be NULL, and the call then
 * only queues its kernels (no host round trip). */
int lsmp_net_apply(lsmp_net_handle net, lsmp_handle h, double band_cells, int *n_band);
/* steps Adam steps (lr, beta 0.9/0.999, eps 1e-8) on minibatches of `batch`
 * pairs drawn with replacement, loss = mean of w (o - y)^2 with
 * w = c^2 / (y^2 + c^2) for weight_c = c > 0 (relative error for |y| >> c)
 * and w = 1 for weight_c = 0; *loss = that mean over the last minibatch. */
int lsmp_net_train(lsmp_net_handle net, int steps, int batch, double lr, unsigned seed, double weight_c,
                   double *loss);
/* Loss and its gradient for the n host samples x[n][128], y[n]; grad has
 * lsmp_net_param_count() entries. */
int lsmp_net_eval_grad(lsmp_net_handle net, const float *x, const float *y, int n, double *loss, float *grad);
/* Learned reinitialisation of the solver's phi: network on |phi| < band_cells dx,
 * then `sweeps` plain pseudo-steps with the sign of the input field.  If the
 * residual mean | |grad phi| - 1 | over |phi| < residual_band dx then exceeds
 * residual_tol (*fallback = 1):
 *   LSMP_FALLBACK_RESTART  -- the input field is reinitialised by the plain
 *                             method instead (its result exactly);
 *   LSMP_FALLBACK_CONTINUE -- plain pseudo-steps continue from the current
 *                             field, `sweeps` at a time, until the residual
 *                             passes, the plain stop (max change <= 1e-6) is
 *                             met or the plain pseudo-time budget is spent;
 *   LSMP_FALLBACK_CONTINUE_RELATIVE -- as CONTINUE, and the continuation also
 *                             stops once a group of `sweeps` pseudo-steps
 *                             lowers the residual by less than 10 %: the
 *                             plain method has then reached what it can
 *                             (thin sheets keep a residual above any fixed
 *                             tolerance).
 * net = NULL runs the same sequence without the network pass (the control
 * that tells the network's share of the result from the sweeps').
 * *residual is the final field's residual, *n_pseudo all pseudo-steps run. */
#define LSMP_FALLBACK_RESTART 0
#define LSMP_FALLBACK_CONTINUE 1
#define LSMP_FALLBACK_CONTINUE_RELATIVE 2
int lsmp_reinit_learned(lsmp_handle h, lsmp_net_handle net, double band_cells, int sweeps, double residual_tol,
                        double residual_band, int fallback_mode, double *residual, int *fallback, int *n_pseudo);
/* The band residual mean | |grad phi| - 1 | over |phi| < band_cells dx, with the
 * Godunov gradient of the current sign field. */
int lsmp_reinit_residual(lsmp_handle h, double band_cells, double *residual);

#ifdef __cplusplus
}
#endif
#endif
