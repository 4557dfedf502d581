/* Test stand-in for liblsmp, for the CPU tests of the Fortran driver's host
 * logic (tests/test_driver_mock.py): every entry point the driver calls
 * appends its name to the file named by LSMP_MOCK_LOG and succeeds.  It
 * computes nothing: the sums report 1, the shape errors 0.  Never linked
 * into the product. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lsmp.h"

struct lsmp_ctx {
    int unused;
};

static struct lsmp_ctx g_ctx;

static void rec(const char *name)
{
    const char *path = getenv("LSMP_MOCK_LOG");
    if (!path) return;
    FILE *f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s\n", name);
    fclose(f);
}

const char *lsmp_error_string(int code) { (void)code; return "mock"; }
int lsmp_init(const lsmp_params *p, lsmp_handle *out) { (void)p; rec("lsmp_init"); *out = &g_ctx; return LSMP_OK; }
void lsmp_finalize(lsmp_handle h) { (void)h; rec("lsmp_finalize"); }
int lsmp_set_advection_scheme(lsmp_handle h, int s) { (void)h; (void)s; rec("lsmp_set_advection_scheme"); return LSMP_OK; }
int lsmp_upload_phi(lsmp_handle h, const double *b) { (void)h; (void)b; rec("lsmp_upload_phi"); return LSMP_OK; }
int lsmp_store_phi0(lsmp_handle h) { (void)h; rec("lsmp_store_phi0"); return LSMP_OK; }
int lsmp_reinit(lsmp_handle h, int btn, double ts, double *err, int *n)
{
    (void)h; (void)ts;
    rec(btn == 0 ? "lsmp_reinit(0)" : "lsmp_reinit(1)");
    if (err) *err = 0.0;
    if (n) *n = 0;
    return LSMP_OK;
}
int lsmp_mass_vol(lsmp_handle h, double *m, double *v) { (void)h; rec("lsmp_mass_vol"); *m = 1.0; *v = 1.0; return LSMP_OK; }
int lsmp_mass_fix(lsmp_handle h, double im, double *m, int *n)
{
    (void)h; (void)im;
    rec("lsmp_mass_fix");
    if (m) *m = 1.0;
    if (n) *n = 1;
    return LSMP_OK;
}
int lsmp_synchronize(lsmp_handle h) { (void)h; rec("lsmp_synchronize"); return LSMP_OK; }
int lsmp_set_velocity_analytic(lsmp_handle h, int f, double t, double t2s)
{
    (void)h; (void)f; (void)t; (void)t2s;
    rec("lsmp_set_velocity_analytic");
    return LSMP_OK;
}
int lsmp_rk3_advect(lsmp_handle h, double dt) { (void)h; (void)dt; rec("lsmp_rk3_advect"); return LSMP_OK; }
int lsmp_get_stats(lsmp_handle h, lsmp_stats *s) { (void)h; memset(s, 0, sizeof *s); return LSMP_OK; }
int lsmp_interface_l1(lsmp_handle h, double *l1) { (void)h; rec("lsmp_interface_l1"); *l1 = 0.0; return LSMP_OK; }
int lsmp_upload_vof(lsmp_handle h, const double *b) { (void)h; (void)b; rec("lsmp_upload_vof"); return LSMP_OK; }
int lsmp_store_vof0(lsmp_handle h) { (void)h; rec("lsmp_store_vof0"); return LSMP_OK; }
int lsmp_store_vof_old(lsmp_handle h) { (void)h; rec("lsmp_store_vof_old"); return LSMP_OK; }
int lsmp_vof_mass_vol(lsmp_handle h, double *m, double *v) { (void)h; rec("lsmp_vof_mass_vol"); *m = 1.0; *v = 1.0; return LSMP_OK; }
int lsmp_set_face_velocity_analytic(lsmp_handle h, int f, double t, double t2s)
{
    (void)h; (void)f; (void)t; (void)t2s;
    rec("lsmp_set_face_velocity_analytic");
    return LSMP_OK;
}
int lsmp_vof_wlic(lsmp_handle h, double dt) { (void)h; (void)dt; rec("lsmp_vof_wlic"); return LSMP_OK; }
int lsmp_clsvof_recon(lsmp_handle h, int *a, int *b)
{
    (void)h;
    rec("lsmp_clsvof_recon");
    if (a) *a = 0;
    if (b) *b = 0;
    return LSMP_OK;
}
int lsmp_vof_interface_l1(lsmp_handle h, double *l1) { (void)h; rec("lsmp_vof_interface_l1"); *l1 = 0.0; return LSMP_OK; }
int lsmp_energy_momentum(lsmp_handle h, double mu_12, int vof_marker, double *out)
{
    (void)h; (void)mu_12; (void)vof_marker;
    rec("lsmp_energy_momentum");
    for (int q = 0; q < 11; ++q) out[q] = 1.0;
    return LSMP_OK;
}
