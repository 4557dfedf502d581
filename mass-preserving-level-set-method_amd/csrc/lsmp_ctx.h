// lsmp_ctx.h -- state of one solver instance, shared by the host translation
// units of the library.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/lsmp.h"
#include "lsmp_internal.h"

struct Pending {
    int phase;            // -1: kernel timing only
    std::string kname;    // empty: phase timing only
    hipEvent_t e0, e1;
};

struct lsmp_ctx {
    lsmp_params p{};
    int dim = 3, nx = 0, ny = 0, nz = 1;
    size_t ncell = 0;
    bool fp32 = false;
    size_t esz = 8;
    hipStream_t st = nullptr;
    // device fields (element type double or float)
    void *phi = nullptr, *phiB = nullptr, *phiC = nullptr;
    void *u = nullptr, *v = nullptr, *w = nullptr;
    void *R = nullptr, *L1 = nullptr, *L2 = nullptr, *S0 = nullptr, *phi0 = nullptr;
    void *dxv = nullptr, *dyv = nullptr, *gr = nullptr, *scr = nullptr;
    void *coef[3] = {nullptr, nullptr, nullptr};
    void *ucoef[3] = {nullptr, nullptr, nullptr};   // UCCD tables (upwind then downwind) per direction
    int scheme = LSMP_SCHEME_CRWENO;
    double *partial = nullptr, *scal = nullptr;
    double *vtab = nullptr;   // per-axis sine tables of the analytic velocity
    unsigned long long *u64 = nullptr;
    unsigned long long *rctl = nullptr;   // pseudo-time loop control block (k_redis_ctl)
    double *hbuf = nullptr;   // pinned readback
    int npartial = 0;
    int band_lines = 1;              // mass correction on band lines: 0 never, 1 where it pays, 2 always
    unsigned char *lflag = nullptr;  // band flags of the x, y, z lines (band_flag_bytes)
    unsigned *lcount = nullptr;      // listed lines per direction
    int *llist[3] = {nullptr, nullptr, nullptr};
    int last_lines[3] = {-1, -1, -1};  // listed lines of the last mass correction
    double nb_cells = 0.0;             // narrow band half-width in cells (0: off)
    unsigned char *tiles = nullptr;    // narrow band: 3D reinitialisation workgroups holding a tube cell
    // 3D reinitialisation, band cells: |grad| side field, the list of band
    // cells and its two lengths, used by turns (allocated on the first btn = 1
    // reinitialisation); rb_count[rb_par] is zero between stages
    void *rb_gr = nullptr;
    unsigned *rb_list = nullptr, *rb_count = nullptr;
    int rb_par = 0;
    int *band_list = nullptr;        // narrow-band cell indices (learned reinitialisation)
    unsigned *band_count = nullptr;
    void *band_val = nullptr;        // the network's value per band entry
    void *mk[2] = {nullptr, nullptr};   // the two markers' fields (two-marker MPLS branch)
    unsigned *mflag = nullptr;          // "inside both markers" flag of the merge
    bool mk_set[2] = {false, false};
    // CLSVOF, allocated on first use: vof%now (and its spare for the passes),
    // vof%old, the kept phi of the reconstruction, vof%tmp2 (the initial VOF of
    // the shape error), face velocities vel%x/y/z,
    // CCD normals, the face velocity tables
    void *vof = nullptr, *vofB = nullptr, *vof_old = nullptr, *vkeep = nullptr, *vof0 = nullptr;
    void *uf = nullptr, *vf = nullptr, *wf = nullptr;
    void *nrm[3] = {nullptr, nullptr, nullptr};
    double *ftab = nullptr;
    // energy_and_momentum, allocated on first use: the nine velocity derivatives
    // and the block sums with the eleven results after them
    void *em = nullptr;
    double *em_part = nullptr;
    bool vof_set = false, vof_old_set = false, vof0_set = false, face_set = false;
    bool vel_set = false, phi0_set = false;
    lsmp_stats stats{};
    bool profiling = false;
    std::map<std::string, std::pair<long long, double>> ktime;
    std::vector<Pending> pend;
    std::vector<hipEvent_t> pool;
};

// Plain reinitialisation (lsmp_host.hip), also used by the learned path.
int lsmp_reinit_run(lsmp_ctx *c, int btn, double timestop, int max_iter, bool init_s0, double *err, int *n);
// The same run queued without its closing read of the control block, when its
// pseudo-steps fit one queued chunk (*launched > 0; else it ran as
// lsmp_reinit_run does and *launched = 0).  While it is pending, the result is
// in c->phiB if (launched - rctl[2]) is odd, else in c->phi;
// lsmp_reinit_finish takes a host copy of rctl[0..3] read after the queued
// work and does what lsmp_reinit_run does after its last read.
int lsmp_reinit_queue(lsmp_ctx *c, int btn, int max_iter, bool init_s0, int *launched, double *err, int *n);
int lsmp_reinit_finish(lsmp_ctx *c, int launched, const unsigned long long *ctl4, double *err, int *n);
lsmp::RedisArgs lsmp_redis_args(const lsmp_ctx *c);
