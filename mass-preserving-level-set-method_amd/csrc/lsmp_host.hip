// lsmp_host.hip -- C ABI of the gfx950 level-set transport library (include/lsmp.h).
//
// Owns the device state of one solver instance, orders the kernel launches of
// each reference call on one HIP stream, and converts between the reference's
// ghosted FP64 host boxes and the interior-only device layout.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/lsmp.h"
#include "lsmp_internal.h"
#include "lsmp_ctx.h"
#include "lsmp_clsvof.h"

using namespace lsmp;

namespace {

constexpr int GHC = 3;

}  // namespace

namespace {

thread_local char g_errbuf[256];

#define LSMP_HIP(call)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            snprintf(g_errbuf, sizeof g_errbuf, "%s: %s", #call, hipGetErrorString(e_));           \
            return LSMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

#define LSMP_LAUNCHED()                                                                            \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess) {                                                                    \
            snprintf(g_errbuf, sizeof g_errbuf, "kernel launch: %s", hipGetErrorString(e_));       \
            return LSMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

hipEvent_t get_event(lsmp_ctx *c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Resolve finished timing records (all of them when wait is true).
void resolve(lsmp_ctx *c, bool wait)
{
    size_t keep = 0;
    for (size_t q = 0; q < c->pend.size(); ++q) {
        Pending &r = c->pend[q];
        if (!wait && hipEventQuery(r.e1) != hipSuccess) {
            c->pend[keep++] = r;
            continue;
        }
        (void)hipEventSynchronize(r.e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.e0, r.e1);
        if (r.phase == 0) c->stats.t_advect_ms += ms;
        else if (r.phase == 1) c->stats.t_reinit_ms += ms;
        else if (r.phase == 2) c->stats.t_massfix_ms += ms;
        if (!r.kname.empty()) {
            auto &kt = c->ktime[r.kname];
            kt.first += 1;
            kt.second += ms;
        }
        c->pool.push_back(r.e0);
        c->pool.push_back(r.e1);
    }
    c->pend.resize(keep);
}

// Record a start event; returns it (or null when not timing).
struct Timer {
    lsmp_ctx *c;
    int phase;
    std::string name;
    hipEvent_t e0 = nullptr;
    Timer(lsmp_ctx *c_, int phase_, const char *name_, bool on) : c(c_), phase(phase_)
    {
        if (!on) return;
        if (name_) name = name_;
        e0 = get_event(c);
        if (e0) (void)hipEventRecord(e0, c->st);
    }
    void stop()
    {
        if (!e0) return;
        hipEvent_t e1 = get_event(c);
        if (!e1) return;
        (void)hipEventRecord(e1, c->st);
        c->pend.push_back(Pending{phase, name, e0, e1});
        e0 = nullptr;
        if (c->pend.size() > 4096) resolve(c, true);
    }
};

// Kernel-level timing: only when profiling is on.
#define KT(ctx, kname, launch)                                                                     \
    do {                                                                                           \
        Timer kt_((ctx), -1, (kname), (ctx)->profiling);                                           \
        launch;                                                                                    \
        LSMP_LAUNCHED();                                                                           \
        kt_.stop();                                                                                \
    } while (0)

int check_params(const lsmp_params *p)
{
    if (!p) return LSMP_E_INVALID;
    if (p->ghc != GHC) return LSMP_E_INVALID;
    if (p->dim != 2 && p->dim != 3) return LSMP_E_INVALID;
    if (p->nx < 4 || p->ny < 4) return LSMP_E_INVALID;
    if (p->dim == 3 && p->nz < 4) return LSMP_E_INVALID;
    if (p->dim == 2 && p->nz != 1) return LSMP_E_INVALID;
    if (!(p->dx > 0) || !(p->dy > 0) || !(p->dz > 0) || !(p->ls_wid > 0) || !(p->rdt > 0))
        return LSMP_E_INVALID;
    if (p->precision != LSMP_FP64 && p->precision != LSMP_FP32) return LSMP_E_INVALID;
    // line indices and per-line scratch offsets are kept in 32-bit ints
    const double lines = (double)p->nx * p->ny * (p->dim == 3 ? p->nz : 1);
    if (lines > 2.0e9) return LSMP_E_INVALID;
    return LSMP_OK;
}

// CCD coefficient tables (ccd_solvers.f90:80-163), factorised once by the
// 2x2 block elimination of TWIN_DEC (matrix_solver.f90:6-45), for a box line
// of nb points with spacing dx.  Output: A, B, AA, BB, each (3, nb), then the
// nb reciprocal pivot determinants.
// kind 0: CCD (init_ccd, ccd_solvers.f90:136-163); 1 / 2: the upwind / downwind
// systems of UCCD (init_uccd, :165-214).  All share the boundary rows of bc
// (:80-134); the tables are factorised by TWIN_DEC (matrix_solver.f90:6-45).
void ccd_tables(int nb, double dx, std::vector<double> &out, int kind = 0)
{
    out.assign((size_t)12 * nb, 0.0);
    double *A = out.data(), *B = A + 3 * nb, *AA = A + 6 * nb, *BB = A + 9 * nb;
    auto M = [&](double *X, int m, int i) -> double & { return X[m + 3 * i]; };
    const int is = 0, ie = nb - 1;
    M(A, 1, is) = 1.0;  M(B, 1, is) = 0.0;  M(AA, 1, is) = 0.0;  M(BB, 1, is) = 1.0;
    M(A, 2, is) = 2.0;  M(B, 2, is) = -dx;  M(AA, 2, is) = -2.5 / dx;  M(BB, 2, is) = 8.5;
    M(A, 1, ie) = 1.0;  M(B, 1, ie) = 0.0;  M(AA, 1, ie) = 0.0;  M(BB, 1, ie) = 1.0;
    M(A, 0, ie) = 2.0;  M(B, 0, ie) = dx;   M(AA, 0, ie) = 2.5 / dx;  M(BB, 0, ie) = 8.5;
    const double A1 = 0.875, B1 = 0.1251282341599089, B2 = -0.2487176584009104, B3 = 0.0001282341599089;
    for (int i = is + 1; i <= ie - 1; ++i) {
        if (kind == 0) {
            M(A, 0, i) = 7.0 / 16.0;  M(A, 1, i) = 1.0;  M(A, 2, i) = 7.0 / 16.0;
            M(B, 0, i) = dx / 16.0;   M(B, 1, i) = 0.0;  M(B, 2, i) = -dx / 16.0;
        } else if (kind == 1) {
            M(A, 0, i) = A1;       M(A, 1, i) = 1.0;      M(A, 2, i) = 0.0;
            M(B, 0, i) = B1 * dx;  M(B, 1, i) = B2 * dx;  M(B, 2, i) = B3 * dx;
        } else {
            M(A, 0, i) = 0.0;       M(A, 1, i) = 1.0;       M(A, 2, i) = A1;
            M(B, 0, i) = -B3 * dx;  M(B, 1, i) = -B2 * dx;  M(B, 2, i) = -B1 * dx;
        }
        M(AA, 0, i) = -(9.0 / 8.0 / dx);  M(AA, 1, i) = 0.0;  M(AA, 2, i) = 9.0 / 8.0 / dx;
        M(BB, 0, i) = -1.0 / 8.0;  M(BB, 1, i) = 1.0;  M(BB, 2, i) = -1.0 / 8.0;
    }
    M(A, 0, is) = 0.0;  M(AA, 0, is) = 0.0;  M(A, 2, ie) = 0.0;  M(AA, 2, ie) = 0.0;
    M(B, 0, is) = 0.0;  M(BB, 0, is) = 0.0;  M(B, 2, ie) = 0.0;  M(BB, 2, ie) = 0.0;
    for (int i = is + 1; i <= ie; ++i) {
        const double den = M(A, 1, i - 1) * M(BB, 1, i - 1) - M(AA, 1, i - 1) * M(B, 1, i - 1);
        double sc1 = -(M(AA, 1, i - 1) * M(B, 0, i) - M(A, 0, i) * M(BB, 1, i - 1)) / den;
        double sc2 = (M(B, 0, i) * M(A, 1, i - 1) - M(B, 1, i - 1) * M(A, 0, i)) / den;
        M(A, 0, i) = sc1;
        M(A, 1, i) = M(A, 1, i) - (sc1 * M(A, 2, i - 1) + sc2 * M(AA, 2, i - 1));
        M(B, 0, i) = sc2;
        M(B, 1, i) = M(B, 1, i) - (sc1 * M(B, 2, i - 1) + sc2 * M(BB, 2, i - 1));
        sc1 = -(M(AA, 1, i - 1) * M(BB, 0, i) - M(AA, 0, i) * M(BB, 1, i - 1)) / den;
        sc2 = (M(BB, 0, i) * M(A, 1, i - 1) - M(B, 1, i - 1) * M(AA, 0, i)) / den;
        M(AA, 0, i) = sc1;
        M(AA, 1, i) = M(AA, 1, i) - (sc1 * M(A, 2, i - 1) + sc2 * M(AA, 2, i - 1));
        M(BB, 0, i) = sc2;
        M(BB, 1, i) = M(BB, 1, i) - (sc1 * M(B, 2, i - 1) + sc2 * M(BB, 2, i - 1));
    }
    // reciprocal of each row's 2x2 pivot determinant, the Den of TWIN_BKS
    // (matrix_solver.f90:61, 71), which depends on the constant tables only
    out.resize((size_t)13 * nb);
    A = out.data(); B = A + 3 * nb; AA = A + 6 * nb; BB = A + 9 * nb;
    double *iden = A + 12 * nb;
    for (int i = 0; i < nb; ++i)
        iden[i] = 1.0 / (M(A, 1, i) * M(BB, 1, i) - M(AA, 1, i) * M(B, 1, i));
}

inline long box_index(const lsmp_ctx *c, int i, int j, int k)   // 0-based interior coords
{
    const long nxb = c->nx + 2 * GHC, nyb = c->ny + 2 * GHC;
    return ((long)(k + GHC) * nyb + (j + GHC)) * nxb + (i + GHC);
}

// host box -> interior (double) vector
void pack_box(const lsmp_ctx *c, const double *box, std::vector<double> &out)
{
    out.resize(c->ncell);
    size_t q = 0;
    for (int k = 0; k < c->nz; ++k)
        for (int j = 0; j < c->ny; ++j) {
            const double *row = box + box_index(c, 0, j, k);
            std::memcpy(&out[q], row, sizeof(double) * c->nx);
            q += c->nx;
        }
}

// interior -> host box with zero-gradient ghosts (branches.f90:283-310)
void unpack_box(const lsmp_ctx *c, const std::vector<double> &in, double *box)
{
    const int nxb = c->nx + 2 * GHC, nyb = c->ny + 2 * GHC, nzb = c->nz + 2 * GHC;
    for (int kb = 0; kb < nzb; ++kb) {
        const int k = std::min(std::max(kb - GHC, 0), c->nz - 1);
        for (int jb = 0; jb < nyb; ++jb) {
            const int j = std::min(std::max(jb - GHC, 0), c->ny - 1);
            const double *src = &in[((size_t)k * c->ny + j) * c->nx];
            double *dst = box + ((long)kb * nyb + jb) * nxb;
            for (int ib = 0; ib < nxb; ++ib) {
                const int i = std::min(std::max(ib - GHC, 0), c->nx - 1);
                dst[ib] = src[i];
            }
        }
    }
}

int upload_field(lsmp_ctx *c, void *dst, const double *box)
{
    std::vector<double> tmp;
    pack_box(c, box, tmp);
    if (c->fp32) {
        std::vector<float> f(tmp.begin(), tmp.end());
        LSMP_HIP(hipMemcpyAsync(dst, f.data(), sizeof(float) * c->ncell, hipMemcpyHostToDevice, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
    } else {
        LSMP_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
    }
    return LSMP_OK;
}

int download_field(lsmp_ctx *c, const void *src, double *box)
{
    std::vector<double> tmp(c->ncell);
    if (c->fp32) {
        std::vector<float> f(c->ncell);
        LSMP_HIP(hipMemcpyAsync(f.data(), src, sizeof(float) * c->ncell, hipMemcpyDeviceToHost, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
        for (size_t q = 0; q < c->ncell; ++q) tmp[q] = f[q];
    } else {
        LSMP_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
    }
    unpack_box(c, tmp, box);
    return LSMP_OK;
}

int alloc(void **p, size_t bytes)
{
    LSMP_HIP(hipMalloc(p, bytes));
    return LSMP_OK;
}

// ------------------------------------------------------------------ phases
// narrow band half-width W (the largest cell size times the cells asked for)
double tube_width(const lsmp_ctx *c)
{
    return c->nb_cells * std::max(c->p.dx, std::max(c->p.dy, c->p.dz));
}

template <typename T> int advect(lsmp_ctx *c, double dt)
{
    T *R = (T *)c->R, *L1 = (T *)c->L1, *L2 = (T *)c->L2, *scr = (T *)c->scr;
    const T *u = (const T *)c->u, *v = (const T *)c->v, *w = (const T *)c->w;
    const T dx = (T)c->p.dx, dy = (T)c->p.dy, dz = (T)c->p.dz, tdt = (T)dt;
    // Narrow band (3D, CRWENO): phi is clamped to [-W, W], only the lines
    // crossing the tube |phi| < W are solved, and the RK stage updates only
    // cells whose x, y and z lines all cross it (every cell with |phi| < W);
    // the other cells keep their (clamped) values.
    const bool nb = c->nb_cells > 0.0 && c->dim == 3 && c->scheme == LSMP_SCHEME_CRWENO;
    int nl[3] = {0, 0, 0};
    TubeMask tm;
    if (nb) {
        KT(c, "tube_lines", (launch_band_lines<T>(c->st, 3, (T *)c->phi, c->nx, c->ny, c->nz, tube_width(c),
                                                  c->lflag, c->lcount, c->llist[0], c->llist[1], c->llist[2], true)));
        unsigned *hc = reinterpret_cast<unsigned *>(c->hbuf);
        LSMP_HIP(hipMemcpyAsync(hc, c->lcount, 3 * sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
        for (int d = 0; d < 3; ++d) nl[d] = (int)hc[d];
        const size_t nfx = (size_t)c->ny * c->nz, nfy = (size_t)c->nx * c->nz;
        tm.fx = c->lflag;
        tm.fy = c->lflag + band_flag_offset(1, nfx, nfy);
        tm.nx = c->nx;
        tm.ny = c->ny;
    }
    for (int stage = 1; stage <= 3; ++stage) {
        T *phi = (T *)c->phi;
        if (c->scheme == LSMP_SCHEME_UCCD) {
            // level_set_rk3_source_ccd: every solver uses dx (ccd_solvers.f90:406-408);
            // the stage result goes to phiB, which then becomes phi
            T *out = (T *)c->phiB;
            const T *cu[3] = {(const T *)c->ucoef[0], (const T *)c->ucoef[1], (const T *)c->ucoef[2]};
            KT(c, "adv_x", (launch_uccd_pass<T>(c->st, 0, 0, stage, phi, out, u, R, L1, L2, cu[0], scr, c->nx,
                                                c->ny, c->nz, dx, tdt)));
            if (c->dim == 3) {
                KT(c, "adv_y", (launch_uccd_pass<T>(c->st, 1, 1, stage, phi, out, v, R, L1, L2, cu[1], scr, c->nx,
                                                    c->ny, c->nz, dx, tdt)));
                KT(c, "adv_z", (launch_uccd_pass<T>(c->st, 2, 2, stage, phi, out, w, R, L1, L2, cu[2], scr, c->nx,
                                                    c->ny, c->nz, dx, tdt)));
            } else {
                KT(c, "adv_y", (launch_uccd_pass<T>(c->st, 1, 2, stage, phi, out, v, R, L1, L2, cu[1], scr, c->nx,
                                                    c->ny, c->nz, dx, tdt)));
            }
            std::swap(c->phi, c->phiB);
            continue;
        }
        if (nb) {
            KT(c, "adv_x", (launch_adv_pass<T>(c->st, 0, 0, stage, phi, u, R, L1, L2, scr, c->nx, c->ny, c->nz, dx,
                                               tdt, c->llist[0], nl[0])));
            KT(c, "adv_y", (launch_adv_pass<T>(c->st, 1, 1, stage, phi, v, R, L1, L2, scr, c->nx, c->ny, c->nz, dy,
                                               tdt, c->llist[1], nl[1])));
            KT(c, "adv_z", (launch_adv_pass<T>(c->st, 2, 2, stage, phi, w, R, L1, L2, scr, c->nx, c->ny, c->nz, dz,
                                               tdt, c->llist[2], nl[2], tm)));
            continue;
        }
        KT(c, "adv_x", (launch_adv_pass<T>(c->st, 0, 0, stage, phi, u, R, L1, L2, scr, c->nx, c->ny, c->nz, dx, tdt)));
        if (c->dim == 3) {
            KT(c, "adv_y", (launch_adv_pass<T>(c->st, 1, 1, stage, phi, v, R, L1, L2, scr, c->nx, c->ny, c->nz, dy, tdt)));
            KT(c, "adv_z", (launch_adv_pass<T>(c->st, 2, 2, stage, phi, w, R, L1, L2, scr, c->nx, c->ny, c->nz, dz, tdt)));
        } else {
            KT(c, "adv_y", (launch_adv_pass<T>(c->st, 1, 2, stage, phi, v, R, L1, L2, scr, c->nx, c->ny, c->nz, dy, tdt)));
        }
    }
    return LSMP_OK;
}

RedisArgs redis_args(const lsmp_ctx *c)
{
    RedisArgs ra;
    ra.nx = c->nx; ra.ny = c->ny; ra.nz = c->nz;
    ra.inv12dx = 1.0 / (12.0 * c->p.dx);
    ra.inv12dy = 1.0 / (12.0 * c->p.dy);
    ra.inv12dz = 1.0 / (12.0 * c->p.dz);
    ra.invdx = 1.0 / c->p.dx;
    ra.invdy = 1.0 / c->p.dy;
    ra.invdz = 1.0 / c->p.dz;
    ra.ls_wid = c->p.ls_wid;
    ra.inv_ls_wid = 1.0 / c->p.ls_wid;
    ra.rdt = c->p.rdt;
    return ra;
}

int reinit_finish(lsmp_ctx *c, int launched, const unsigned long long *hc, double *red_error, int *n_pseudo)
{
    // each queued step that the stop flag skipped still exchanged phi and phiB
    const int iter = (int)hc[2];
    if ((launched - iter) % 2) std::swap(c->phi, c->phiB);
    double error;
    std::memcpy(&error, &hc[3], sizeof error);
    if (std::isnan(error)) return LSMP_E_NAN;
    c->stats.last_redis_iters = iter;
    c->stats.last_red_error = error;
    if (red_error) *red_error = error;
    if (n_pseudo) *n_pseudo = iter;
    return LSMP_OK;
}

// queued (non-null): queue the steps without the closing read when they fit
// one chunk, *queued = the steps launched (0 when the run went the usual way)
template <typename T>
int reinit(lsmp_ctx *c, int btn, double timestop, int max_iter, bool init_s0, double *red_error, int *n_pseudo,
           int *queued = nullptr)
{
    const RedisArgs ra = redis_args(c);
    const size_t n = c->ncell;
    T *S0 = (T *)c->S0;
    // level_set_redis_init / level_set_redis_stable (redis.f90:81-174)
    if (!init_s0) {
        // S0 was set by the caller (the learned path keeps the sign of the
        // field before the network's correction)
    } else if (btn == 1) {
        KT(c, "sign_init", (launch_sign_init<T>(c->st, (const T *)c->phi, S0, n, (T)c->p.ls_wid, 0)));
    } else {
        KT(c, "sign_init", (launch_sign_init<T>(c->st, (const T *)c->phi, S0, n, (T)c->p.ls_wid, 1)));
        launch_zero_u64(c->st, c->u64 + 1, 1);
        KT(c, "grad_max", (launch_grad_max<T>(c->st, c->dim, (const T *)c->phi, S0, ra, c->u64 + 1)));
        KT(c, "scale_max", (launch_scale_by_max<T>(c->st, (T *)c->phi, n, c->u64 + 1)));
    }
    if (timestop < 0.0) {
        if (btn == 0) {
            double m = c->p.xend - c->p.xstart;
            m = std::max(m, c->p.yend - c->p.ystart);
            m = std::max(m, c->p.zend - c->p.zstart);
            timestop = 1.5 * m;
        } else {
            timestop = 2.0 * c->p.ls_wid;
        }
    }
    // pseudo-steps until time > timestop, accumulated as the reference does
    int budget = 0;
    for (double time = 0.0; budget < 1000000;) {
        ++budget;
        time = time + c->p.rdt;
        if (time > timestop) break;
    }
    if (max_iter > 0 && budget > max_iter) budget = max_iter;
    // the max |dphi| <= 1e-6 stop is taken on the device (k_redis_ctl); steps
    // are queued CHUNK at a time and the host reads the control block per chunk
    constexpr int CHUNK = 16;
    unsigned long long *ctl = c->rctl;
    // narrow band (3D): workgroups with no cell |phi| < W keep their cells
    const unsigned char *tiles = nullptr;
    if (c->nb_cells > 0.0 && c->dim == 3) {
        KT(c, "tube_tiles", (launch_tube_tiles<T>(c->st, (const T *)c->phi, c->nx, c->ny, c->nz, tube_width(c),
                                                  c->tiles)));
        tiles = c->tiles;
    }
    RedisBand<T> bd;
    if (c->dim == 3 && btn == 1) {
        if (!c->rb_list) {
            if (alloc(&c->rb_gr, c->ncell * c->esz) || alloc((void **)&c->rb_list, c->ncell * sizeof(unsigned)) ||
                alloc((void **)&c->rb_count, 2 * sizeof(unsigned)))
                return LSMP_E_HIP;
            launch_zero_bytes(c->st, (unsigned char *)c->rb_count, 2 * sizeof(unsigned));
            c->rb_par = 0;
        }
        bd.gr = (T *)c->rb_gr;
        bd.list = c->rb_list;
    }
    launch_redis_ctl(c->st, ctl, 0);
    int launched = 0;
    unsigned long long hc[4] = {0, 0, 0, 0};
    const bool defer = queued && budget <= CHUNK;
    if (queued) *queued = 0;
    while (launched < budget) {
        const int k = std::min(CHUNK, budget - launched);
        for (int q = 0; q < k; ++q) {
            T *cur = (T *)c->phi, *b = (T *)c->phiB, *cc = (T *)c->phiC;
            if (c->dim == 3) {
                // stages 1..3: every cell (lambda = 0), then the band cells (lambda)
                const T *in[3] = {cur, b, cc};
                T *outp[3] = {b, cc, b};
                for (int sg = 0; sg < 3; ++sg) {
                    if (btn == 1) {
                        bd.count = c->rb_count + c->rb_par;
                        bd.next = c->rb_count + (1 - c->rb_par);
                    }
                    KT(c, "redis_stage",
                       (launch_redis3d<T>(c->st, sg + 1, btn, in[sg], outp[sg], cur, S0, ra, ctl, tiles, bd, false)));
                    if (btn == 1) {
                        KT(c, "redis_band",
                           (launch_redis3d<T>(c->st, sg + 1, btn, in[sg], outp[sg], cur, S0, ra, ctl, tiles, bd, true)));
                        c->rb_par ^= 1;
                    }
                }
            } else {
                KT(c, "redis_stage", (launch_redis_stage2d<T>(c->st, 1, btn, cur, b, cur, S0, ra, ctl)));
                KT(c, "redis_stage", (launch_redis_stage2d<T>(c->st, 2, btn, b, cc, cur, S0, ra, ctl)));
                KT(c, "redis_stage", (launch_redis_stage2d<T>(c->st, 3, btn, cc, b, cur, S0, ra, ctl)));
            }
            launch_redis_ctl(c->st, ctl, 2);
            // the new phi is in b; the old one becomes a spare
            c->phiB = c->phi;
            c->phi = b;
        }
        launched += k;
        if (defer) {
            *queued = launched;
            return LSMP_OK;
        }
        LSMP_HIP(hipMemcpyAsync(c->hbuf, ctl, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
        std::memcpy(hc, c->hbuf, sizeof hc);
        if (hc[1]) break;
    }
    return reinit_finish(c, launched, hc, red_error, n_pseudo);
}

CcdArgs ccd_args(const lsmp_ctx *c)
{
    CcdArgs ca;
    ca.dx = c->p.dx; ca.dy = c->p.dy; ca.dz = c->p.dz;
    ca.dv = c->p.dx * c->p.dy * c->p.dz;
    ca.ls_wid = c->p.ls_wid;
    ca.rho_12 = c->p.rho_12;
    ca.inverse = c->p.inverse;
    return ca;
}

template <typename T> int mass_fix(lsmp_ctx *c, double imass, double *mass_out, int *n_iter)
{
    const CcdArgs ca = ccd_args(c);
    T *phi = (T *)c->phi;
    int iter, nblk = 0, rc = LSMP_OK;
    double mass = 0.0, prev_rel = 0.0;
    // The correction changes phi only where delta != 0, so no cell enters the
    // band during the iterations: the lines crossing it at the start cover
    // every later iteration.
    const int *lst[3] = {nullptr, nullptr, nullptr};
    int nlst[3] = {0, 0, 0};
    const int dfin = c->dim == 3 ? 2 : 1;   // direction of the final pass (its sums)
    for (int d = 0; d < 3; ++d) c->last_lines[d] = -1;
    int nfixed = 0;                         // partials of the unlisted lines, constant over the iterations
    // Band lines pay where solving every line would take the one-lane kernel
    // (more than PART_MAX_LINES_STRIDED lines in a direction) and the listed
    // lines fit the 32-lane one: 3D grids from about 65^3 with lines of at
    // most 540 cells.  Elsewhere every line already goes through the 32-lane
    // solve, or the one-lane solve of long lines gains nothing from fewer lanes.
    const int ns[3] = {c->nx, c->ny, c->nz};
    bool pays = c->dim == 3;
    size_t maxl = 0;
    for (int d = 0; d < c->dim; ++d) {
        pays = pays && ns[d] + 4 <= 32 * KCC;
        maxl = std::max(maxl, c->ncell / (size_t)ns[d]);
    }
    pays = pays && maxl > (size_t)PART_MAX_LINES_STRIDED;
    if (c->band_lines == 2 || (c->band_lines == 1 && pays)) {
        KT(c, "band_lines", (launch_band_lines<T>(c->st, c->dim, phi, c->nx, c->ny, c->nz,
                                                  c->p.ls_wid * (1.0 + 1.0e-6), c->lflag, c->lcount, c->llist[0],
                                                  c->llist[1], c->llist[2], false)));
        unsigned *hc = reinterpret_cast<unsigned *>(c->hbuf);
        LSMP_HIP(hipMemcpyAsync(hc, c->lcount, (size_t)c->dim * sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
        for (int d = 0; d < c->dim; ++d) {
            lst[d] = c->llist[d];
            nlst[d] = (int)hc[d];
        }
        for (int d = 0; d < 3; ++d) c->last_lines[d] = nlst[d];
        const unsigned char *ffin =
            c->lflag + band_flag_offset(dfin, (size_t)c->ny * c->nz, (size_t)c->nx * c->nz);
        // the unlisted lines have delta = 0 on every cell, so the correction
        // leaves them as they are: their mass and volume are summed once
        KT(c, "line_hsums", (launch_line_hsums<T>(c->st, dfin, phi, ffin, c->nx, c->ny, c->nz, ca, c->partial,
                                                  &nfixed)));
    }
    double *psolve = c->partial + (size_t)3 * nfixed;
    for (iter = 1; iter <= 15; ++iter) {
        KT(c, "ccd_x", (launch_ccd_pass<T>(c->st, 0, c->dim, false, phi, (const T *)c->coef[0], (T *)c->scr,
                                           (T *)c->dxv, nullptr, nullptr, nullptr, nullptr, c->nx, c->ny, c->nz,
                                           ca, nullptr, lst[0], nlst[0])));
        if (c->dim == 3) {
            KT(c, "ccd_y", (launch_ccd_pass<T>(c->st, 1, 3, false, phi, (const T *)c->coef[1], (T *)c->scr,
                                               (T *)c->dyv, nullptr, nullptr, nullptr, nullptr, c->nx, c->ny, c->nz,
                                               ca, nullptr, lst[1], nlst[1])));
            KT(c, "ccd_z", (launch_ccd_pass<T>(c->st, 2, 3, true, phi, (const T *)c->coef[2], (T *)c->scr, nullptr,
                                               (const T *)c->dxv, (const T *)c->dyv, (T *)c->gr, psolve, c->nx,
                                               c->ny, c->nz, ca, &nblk, lst[2], nlst[2])));
        } else {
            KT(c, "ccd_y", (launch_ccd_pass<T>(c->st, 1, 2, true, phi, (const T *)c->coef[1], (T *)c->scr, nullptr,
                                               (const T *)c->dxv, nullptr, (T *)c->gr, psolve, c->nx, c->ny,
                                               c->nz, ca, &nblk, lst[1], nlst[1])));
        }
        nblk += nfixed;
        if (nblk > c->npartial) return LSMP_E_INVALID;
        KT(c, "finalize", (launch_finalize3(c->st, c->partial, nblk, imass, c->scal)));
        if (lst[dfin])
            KT(c, "mass_update", (launch_mass_update_lines<T>(c->st, dfin, phi, (const T *)c->gr, (T)c->p.ls_wid,
                                                              c->scal, lst[dfin], nlst[dfin], c->nx, c->ny, c->nz)));
        else
            KT(c, "mass_update",
               (launch_mass_update<T>(c->st, phi, (const T *)c->gr, c->ncell, (T)c->p.ls_wid, c->scal)));
        LSMP_HIP(hipMemcpyAsync(c->hbuf, c->scal, 5 * sizeof(double), hipMemcpyDeviceToHost, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
        mass = c->hbuf[1];
        if (c->hbuf[4] != 0.0) { rc = LSMP_E_EMPTY_BAND; break; }
        if (std::isnan(mass) || std::isnan(c->hbuf[3])) { rc = LSMP_E_NAN; break; }
        const double rel = std::fabs(imass - mass) / imass;
        if (rel < 1.0e-14) break;   // the reference's exit test (mass_preserving_level_set.f90:70)
        if (c->fp32) {
            // FP32 bulk phi resolves the mass only to ~1e-10..1e-8 relative; stop
            // once the correction no longer halves the defect
            if (rel < 1.0e-10 || (iter > 1 && rel >= 0.5 * prev_rel)) break;
        }
        prev_rel = rel;
    }
    if (iter > 15) iter = 15;
    c->stats.last_mass_iters = iter;
    if (mass_out) *mass_out = mass;
    if (n_iter) *n_iter = iter;
    return rc;
}

int reduce_partials(lsmp_ctx *c, int nblk, double *out3)
{
    launch_finalize3(c->st, c->partial, nblk, 0.0, c->scal);
    LSMP_LAUNCHED();
    LSMP_HIP(hipMemcpyAsync(c->hbuf, c->scal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    LSMP_HIP(hipStreamSynchronize(c->st));
    out3[0] = c->hbuf[0];
    out3[1] = c->hbuf[1];
    out3[2] = c->hbuf[2];
    return LSMP_OK;
}

int reduce_blocks(const lsmp_ctx *c)
{
    size_t b = (c->ncell + 255) / 256;
    if (b > 1024) b = 1024;
    if ((int)b > c->npartial) b = c->npartial;
    return (int)b;
}

template <typename T> int mass_vol(lsmp_ctx *c, double *mass, double *vol)
{
    const int nb = reduce_blocks(c);
    KT(c, "mass_vol", (launch_mass_vol<T>(c->st, (const T *)c->phi, c->ncell, ccd_args(c), c->partial, nb)));
    double r[3];
    int rc = reduce_partials(c, nb, r);
    if (rc) return rc;
    if (vol) *vol = r[0];
    if (mass) *mass = r[1];
    return LSMP_OK;
}

template <typename T> int interface_l1(lsmp_ctx *c, double *l1)
{
    const int nb = reduce_blocks(c);
    KT(c, "l1", (launch_l1<T>(c->st, (const T *)c->phi, (const T *)c->phi0, c->ncell, c->p.ls_wid, c->partial, nb)));
    double r[3];
    int rc = reduce_partials(c, nb, r);
    if (rc) return rc;
    *l1 = r[0] / (double)c->nx / (double)c->ny / (double)c->nz;
    return LSMP_OK;
}

// energy_and_momentum (output.f90:250-397): the nine first CCD derivatives of
// the velocity along every line, then the eleven integrals (lsmp_energy.hip)
template <typename T> int energy_momentum(lsmp_ctx *c, double mu_12, bool vof_marker, double *out)
{
    if (!c->vel_set || (vof_marker && !c->vof_set)) return LSMP_E_STATE;
    constexpr int NBMAX = 1024;
    if (!c->em) {
        if (alloc(&c->em, (size_t)9 * c->ncell * c->esz) ||
            alloc((void **)&c->em_part, ((size_t)ENERGY_SUMS * NBMAX + ENERGY_SUMS) * sizeof(double)))
            return LSMP_E_HIP;
    }
    const CcdArgs ca = ccd_args(c);
    const T *vel[3] = {(const T *)c->u, (const T *)c->v, (const T *)c->w};
    T *ten = (T *)c->em;
    static const char *const nm[3] = {"em_ccd_x", "em_ccd_y", "em_ccd_z"};
    for (int f = 0; f < c->dim; ++f)
        for (int d = 0; d < c->dim; ++d)
            KT(c, nm[d], (launch_ccd_pass<T>(c->st, d, c->dim, false, vel[f], (const T *)c->coef[d], (T *)c->scr,
                                             ten + (size_t)(3 * f + d) * c->ncell, nullptr, nullptr, nullptr, nullptr,
                                             c->nx, c->ny, c->nz, ca, nullptr)));
    EnergyArgs ea;
    ea.nx = c->nx; ea.ny = c->ny; ea.nz = c->nz; ea.inverse = c->p.inverse;
    ea.dx = c->p.dx; ea.dv = c->p.dx * c->p.dy * c->p.dz; ea.ls_wid = c->p.ls_wid;
    ea.rho_12 = c->p.rho_12; ea.mu_12 = mu_12;
    ea.xstart = c->p.xstart; ea.ystart = c->p.ystart; ea.zstart = c->p.zstart;
    ea.dxc = c->p.dx; ea.dyc = c->p.dy; ea.dzc = c->p.dz;
    const int nb = (int)std::min<size_t>(NBMAX, (c->ncell + 255) / 256);
    KT(c, "energy", (launch_energy_sums<T>(c->st, c->dim, vof_marker, (const T *)c->phi, (const T *)c->vof, vel[0],
                                           vel[1], vel[2], ten, ea, c->em_part, nb)));
    double *res = c->em_part + (size_t)ENERGY_SUMS * NBMAX;
    KT(c, "energy_final", (launch_energy_finalize(c->st, c->em_part, nb, c->p.dx, res)));
    LSMP_HIP(hipMemcpyAsync(c->hbuf, res, ENERGY_SUMS * sizeof(double), hipMemcpyDeviceToHost, c->st));
    LSMP_HIP(hipStreamSynchronize(c->st));
    for (int q = 0; q < ENERGY_SUMS; ++q) out[q] = c->hbuf[q];
    return LSMP_OK;
}

template <typename T> int velocity_analytic(lsmp_ctx *c, int field, double t, double t2s)
{
    const double pi = std::acos(-1.0);
    VelArgs va;
    va.nx = c->nx; va.ny = c->ny; va.nz = c->nz; va.field = field;
    va.xstart = c->p.xstart; va.ystart = c->p.ystart; va.zstart = c->p.zstart;
    va.dx = c->p.dx; va.dy = c->p.dy; va.dz = c->p.dz;
    va.ct = std::cos(pi * t / t2s);
    va.omega = 2.0 * pi / t2s;
    KT(c, "velocity", (launch_velocity<T>(c->st, (T *)c->u, (T *)c->v, (T *)c->w, va, c->vtab)));
    return LSMP_OK;
}

// ---------------------------------------------------------------- CLSVOF
// device fields of the CLSVOF path, allocated (and zeroed) on first use
int ensure_clsvof(lsmp_ctx *c)
{
    const size_t b = c->esz * c->ncell;
    void **f[] = {&c->vof,  &c->vofB, &c->vof_old, &c->vkeep,  &c->vof0,  &c->uf,
                  &c->vf,   &c->wf,   &c->nrm[0],  &c->nrm[1], &c->nrm[2]};
    for (void **p : f)
        if (!*p) {
            LSMP_HIP(hipMalloc(p, b));
            LSMP_HIP(hipMemsetAsync(*p, 0, b, c->st));
        }
    if (!c->ftab) {
        LSMP_HIP(hipMalloc((void **)&c->ftab, sizeof(double) * 3 * (size_t)(c->nx + c->ny + c->nz)));
    }
    LSMP_HIP(hipStreamSynchronize(c->st));
    return LSMP_OK;
}

// face velocities vel%x/y/z of a prescribed field at time t (velocity.f90)
template <typename T> int face_velocity_analytic(lsmp_ctx *c, int field, double t, double t2s)
{
    const double pi = std::acos(-1.0);
    FaceVelArgs va;
    va.nx = c->nx; va.ny = c->ny; va.nz = c->nz; va.field = field;
    va.xstart = c->p.xstart; va.ystart = c->p.ystart; va.zstart = c->p.zstart;
    va.dx = c->p.dx; va.dy = c->p.dy; va.dz = c->p.dz;
    va.ct = std::cos(pi * t / t2s);
    va.omega = 2.0 * pi / t2s;
    KT(c, "face_velocity", (launch_face_velocity<T>(c->st, (T *)c->uf, (T *)c->vf, (T *)c->wf, va, c->p.xend,
                                                     c->p.yend, c->p.zend, c->ftab)));
    return LSMP_OK;
}

// vof_wlic_solver (clsvof.f90:1-273): CCD normals of phi (surface_norms), then
// one WLIC-THINC pass per direction, vof -> vofB -> vof ...
template <typename T> int vof_wlic(lsmp_ctx *c, double dt)
{
    const CcdArgs ca = ccd_args(c);
    static const char *const nname[3] = {"norm_x", "norm_y", "norm_z"};
    // the fluxes read the normals of cut cells only: the CCD solves run on the
    // lines that hold a cell a step can find cut (the same solves, fewer
    // lines; at 256^3 0.29 ms against 0.46 ms for every line)
    const int *lst[3] = {nullptr, nullptr, nullptr};
    int nlst[3] = {0, 0, 0};
    {
        KT(c, "cut_lines", (launch_cut_lines<T>(c->st, c->dim, (const T *)c->vof, c->nx, c->ny, c->nz, c->lflag,
                                                c->lcount, c->llist[0], c->llist[1], c->llist[2])));
        unsigned *hc = reinterpret_cast<unsigned *>(c->hbuf);
        LSMP_HIP(hipMemcpyAsync(hc, c->lcount, (size_t)c->dim * sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        LSMP_HIP(hipStreamSynchronize(c->st));
        for (int d = 0; d < c->dim; ++d) {
            lst[d] = c->llist[d];
            nlst[d] = (int)hc[d];
        }
    }
    for (int d = 0; d < c->dim; ++d)
        KT(c, nname[d], (launch_ccd_pass<T>(c->st, d, c->dim, false, (const T *)c->phi, (const T *)c->coef[d],
                                            (T *)c->scr, (T *)c->nrm[d], nullptr, nullptr, nullptr, nullptr, c->nx,
                                            c->ny, c->nz, ca, nullptr, lst[d], nlst[d])));
    const T *vel[3] = {(const T *)c->uf, (const T *)c->vf, (const T *)c->wf};
    const double dd[3] = {c->p.dx, c->p.dy, c->p.dz};
    const T *n2 = c->dim == 3 ? (const T *)c->nrm[2] : nullptr;
    T *in = (T *)c->vof, *out = (T *)c->vofB;
    for (int d = 0; d < c->dim; ++d) {
        KT(c, "wlic", (launch_wlic_pass<T>(c->st, d, d == c->dim - 1, in, out, (const T *)c->vof_old, vel[d],
                                           (const T *)c->nrm[0], (const T *)c->nrm[1], n2, c->nx, c->ny, c->nz, dt,
                                           dd[d])));
        std::swap(in, out);
    }
    c->vof = in;
    c->vofB = out;
    return LSMP_OK;
}

// clsvof_recon (clsvof.f90:275-324): phi <- 2 vof - 1, reinitialised with
// btn = 0 to T = 5 ls_wid, the kept phi where |phi_kept| > 1.5 ls_wid, then
// the btn = 1 reinitialisation
template <typename T> int clsvof_recon(lsmp_ctx *c, int *n0, int *n1)
{
    KT(c, "recon_start", (launch_recon_start<T>(c->st, (T *)c->phi, (const T *)c->vof, (T *)c->vkeep, c->ncell)));
    double err = 0.0;
    int rc = reinit<T>(c, 0, 5.0 * c->p.ls_wid, 0, true, &err, n0);
    if (rc) return rc;
    KT(c, "recon_keep",
       (launch_recon_keep<T>(c->st, (T *)c->phi, (const T *)c->vkeep, c->ncell, 1.5 * c->p.ls_wid)));
    return reinit<T>(c, 1, -1.0, 0, true, &err, n1);
}

// the VOF half of ls_mv (l1 == 0: *a = volume, *b = mass) or the VOF shape
// error of calculate_interface_loss (l1 != 0: *a = sum |vof0 - vof| / cells)
template <typename T> int vof_sums(lsmp_ctx *c, int l1, double *a, double *b)
{
    const int nb = reduce_blocks(c);
    KT(c, l1 ? "vof_l1" : "vof_mass_vol",
       (launch_vof_sums<T>(c->st, (const T *)c->vof, (const T *)c->vof0, c->ncell, c->p.dx * c->p.dy * c->p.dz,
                           c->p.rho_12, c->p.inverse, l1, c->partial, nb)));
    double r[3];
    int rc = reduce_partials(c, nb, r);
    if (rc) return rc;
    if (l1) {
        *a = r[0] / (double)c->nx / (double)c->ny / (double)c->nz;
    } else {
        *a = r[0];
        *b = r[1];
    }
    return LSMP_OK;
}

void free_all(lsmp_ctx *c)
{
    void *bufs[] = {c->rctl, c->ucoef[0], c->ucoef[1], c->ucoef[2], c->vtab, c->phi, c->phiB, c->phiC, c->u, c->v, c->w, c->R, c->L1, c->L2, c->S0, c->phi0,
                    c->dxv, c->dyv, c->gr, c->scr, c->coef[0], c->coef[1], c->coef[2], c->partial, c->scal, c->u64, c->band_list, c->band_count, c->band_val,
                    c->mk[0], c->mk[1], c->mflag, c->lflag, c->lcount, c->llist[0], c->llist[1], c->llist[2], c->tiles,
                    c->rb_gr, c->rb_list, c->rb_count,
                    c->vof, c->vofB, c->vof_old, c->vkeep, c->vof0, c->uf, c->vf, c->wf, c->nrm[0], c->nrm[1], c->nrm[2], c->ftab, c->em, c->em_part};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->hbuf) (void)hipHostFree(c->hbuf);
    for (auto &r : c->pend) { (void)hipEventDestroy(r.e0); (void)hipEventDestroy(r.e1); }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    if (c->st) (void)hipStreamDestroy(c->st);
}

}  // namespace

// Plain reinitialisation, shared with the learned path: max_iter > 0 caps the
// pseudo-steps, init_s0 = false keeps the sign field already in S0.
int lsmp_reinit_run(lsmp_ctx *c, int btn, double timestop, int max_iter, bool init_s0, double *err, int *n)
{
    return c->fp32 ? reinit<float>(c, btn, timestop, max_iter, init_s0, err, n)
                   : reinit<double>(c, btn, timestop, max_iter, init_s0, err, n);
}

int lsmp_reinit_queue(lsmp_ctx *c, int btn, int max_iter, bool init_s0, int *launched, double *err, int *n)
{
    return c->fp32 ? reinit<float>(c, btn, -1.0, max_iter, init_s0, err, n, launched)
                   : reinit<double>(c, btn, -1.0, max_iter, init_s0, err, n, launched);
}

int lsmp_reinit_finish(lsmp_ctx *c, int launched, const unsigned long long *ctl4, double *err, int *n)
{
    return reinit_finish(c, launched, ctl4, err, n);
}

RedisArgs lsmp_redis_args(const lsmp_ctx *c) { return redis_args(c); }

// ====================================================================== ABI
extern "C" {

const char *lsmp_error_string(int code)
{
    switch (code) {
    case LSMP_OK: return "ok";
    case LSMP_E_INVALID: return "invalid parameter or size";
    case LSMP_E_HIP: return g_errbuf[0] ? g_errbuf : "HIP runtime error";
    case LSMP_E_NAN: return "non-finite value in the level-set field";
    case LSMP_E_EMPTY_BAND: return "mass correction: no cell inside the interface band";
    case LSMP_E_STATE: return "call out of order (velocity, phi0 or marker not set)";
    default: return "unknown error";
    }
}

int lsmp_init(const lsmp_params *params, lsmp_handle *out)
{
    if (!out) return LSMP_E_INVALID;
    *out = nullptr;
    int rc = check_params(params);
    if (rc) return rc;
    lsmp_ctx *c = new lsmp_ctx();
    c->p = *params;
    c->dim = params->dim;
    c->nx = params->nx;
    c->ny = params->ny;
    c->nz = params->dim == 3 ? params->nz : 1;
    c->ncell = (size_t)c->nx * c->ny * c->nz;
    c->fp32 = params->precision == LSMP_FP32;
    c->esz = c->fp32 ? sizeof(float) : sizeof(double);
    auto fail = [&](int code) {
        free_all(c);
        delete c;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) return fail(LSMP_E_HIP);
    const size_t fb = c->ncell * c->esz;
    void **fields[] = {&c->phi, &c->phiB, &c->phiC, &c->u, &c->v, &c->w, &c->R, &c->L1, &c->L2,
                       &c->S0, &c->phi0, &c->dxv, &c->dyv, &c->gr};
    for (void **f : fields)
        if (alloc(f, fb)) return fail(LSMP_E_HIP);
    // scratch: CRWENO (c', d') for n-2 rows of 2 systems per line (x: 64
    // slots per 32 lines), or for n-1-n/2 rows of 4 half systems per line
    // (y, z: 64 slots per 16 lines), or CCD (S, SS) for n+5 rows per line,
    // whichever direction needs most
    size_t sc = 0;
    const int ns[3] = {c->nx, c->ny, c->nz};
    for (int d = 0; d < c->dim; ++d) {
        const size_t nlines = c->ncell / (size_t)ns[d];
        sc = std::max(sc, (size_t)2 * (size_t)(ns[d] - 2) * ((nlines + 31) / 32 * 64));
        sc = std::max(sc, (size_t)2 * (size_t)(ns[d] - 1 - ns[d] / 2) * ((nlines + 15) / 16 * 64));
        // CCD / UCCD checkpoints per line and CSEG-row segment, lanes padded to 256
        const size_t lanes = (nlines + 255) / 256 * 256, segs = (size_t)(ns[d] + 4) / CSEG + 1;
        sc = std::max(sc, (size_t)4 * segs * lanes);   // UCCD keeps 4 (CCD 2)
    }
    if (alloc(&c->scr, sc * c->esz)) return fail(LSMP_E_HIP);
    if (alloc((void **)&c->vtab, (size_t)2 * (c->nx + c->ny + c->nz) * sizeof(double))) return fail(LSMP_E_HIP);
    if (alloc((void **)&c->rctl, CTL_WORDS * sizeof(unsigned long long))) return fail(LSMP_E_HIP);
    for (int d = 0; d < c->dim; ++d) {
        std::vector<double> tab;
        ccd_tables(ns[d] + 2 * GHC, params->dx, tab);   // all three solvers use dx (ccd_solvers.f90:406-408)
        if (alloc(&c->coef[d], tab.size() * c->esz)) return fail(LSMP_E_HIP);
        if (c->fp32) {
            std::vector<float> tf(tab.begin(), tab.end());
            if (hipMemcpy(c->coef[d], tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                return fail(LSMP_E_HIP);
        } else if (hipMemcpy(c->coef[d], tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) !=
                   hipSuccess) {
            return fail(LSMP_E_HIP);
        }
    }
    // partial sums: one slot per block of the last CCD pass (a solve with up to
    // 8 lines per block, then the unlisted lines' sums) or reduction kernel
    size_t maxlines = 0;
    for (int d = 1; d < c->dim; ++d) maxlines = std::max(maxlines, c->ncell / (size_t)ns[d]);
    c->npartial = (int)std::max<size_t>(1024, (maxlines + 7) / 8 + LINE_SEGMENTS * ((maxlines + 255) / 256));
    if (alloc((void **)&c->partial, (size_t)3 * c->npartial * sizeof(double))) return fail(LSMP_E_HIP);
    if (alloc((void **)&c->scal, 8 * sizeof(double))) return fail(LSMP_E_HIP);
    if (alloc((void **)&c->u64, 4 * sizeof(unsigned long long))) return fail(LSMP_E_HIP);
    if (hipHostMalloc((void **)&c->hbuf, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(LSMP_E_HIP);
    // band lines of the mass correction: flags of every x, y, z line, lists of the x and y lines
    const size_t nfx = (size_t)c->ny * c->nz, nfy = (size_t)c->nx * c->nz, nfz = (size_t)c->nx * c->ny;
    if (alloc((void **)&c->lflag, band_flag_bytes(nfx, nfy, nfz)) ||
        alloc((void **)&c->lcount, 4 * sizeof(unsigned)) || alloc((void **)&c->llist[0], nfx * sizeof(int)) ||
        alloc((void **)&c->llist[1], nfy * sizeof(int)) || alloc((void **)&c->llist[2], nfz * sizeof(int)))
        return fail(LSMP_E_HIP);
    if (c->dim == 3 &&
        alloc((void **)&c->tiles,
              (size_t)((c->nx + R3X - 1) / R3X) * ((c->ny + R3Y - 1) / R3Y) * ((c->nz + RKC_NB - 1) / RKC_NB)))
        return fail(LSMP_E_HIP);
    // velocity starts at zero so that an advect before set_velocity is defined;
    // the normals start at zero because a mass correction on band lines leaves
    // the other lines' values as they were (finite, and multiplied by delta = 0)
    if (hipMemsetAsync(c->u, 0, fb, c->st) != hipSuccess || hipMemsetAsync(c->v, 0, fb, c->st) != hipSuccess ||
        hipMemsetAsync(c->w, 0, fb, c->st) != hipSuccess || hipMemsetAsync(c->dxv, 0, fb, c->st) != hipSuccess ||
        hipMemsetAsync(c->dyv, 0, fb, c->st) != hipSuccess || hipMemsetAsync(c->gr, 0, fb, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
        return fail(LSMP_E_HIP);
    *out = c;
    return LSMP_OK;
}

void lsmp_finalize(lsmp_handle h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->st);
    free_all(h);
    delete h;
}

int lsmp_upload_phi(lsmp_handle h, const double *phi_box)
{
    if (!h || !phi_box) return LSMP_E_INVALID;
    return upload_field(h, h->phi, phi_box);
}

int lsmp_download_phi(lsmp_handle h, double *phi_box)
{
    if (!h || !phi_box) return LSMP_E_INVALID;
    return download_field(h, h->phi, phi_box);
}

int lsmp_store_phi0(lsmp_handle h)
{
    if (!h) return LSMP_E_INVALID;
    LSMP_HIP(hipMemcpyAsync(h->phi0, h->phi, h->ncell * h->esz, hipMemcpyDeviceToDevice, h->st));
    LSMP_HIP(hipStreamSynchronize(h->st));
    h->phi0_set = true;
    return LSMP_OK;
}

// ---- two-marker MPLS branch (interface_solver.f90:34-121) ----

int lsmp_set_marker(lsmp_handle h, int mid, const double *box)
{
    if (!h || !box || (mid != 1 && mid != 2)) return LSMP_E_INVALID;
    if (!h->mk[0]) {   // both markers and the merge flag, all or nothing, on first use
        void *a = nullptr, *b = nullptr, *f = nullptr;
        if (alloc(&a, h->ncell * h->esz) || alloc(&b, h->ncell * h->esz) || alloc(&f, sizeof(unsigned))) {
            for (void *q : {a, b, f})
                if (q) (void)hipFree(q);
            return LSMP_E_HIP;
        }
        h->mk[0] = a;
        h->mk[1] = b;
        h->mflag = (unsigned *)f;
    }
    const int rc = upload_field(h, h->mk[mid - 1], box);
    if (rc == LSMP_OK) h->mk_set[mid - 1] = true;
    return rc;
}

int lsmp_get_marker(lsmp_handle h, int mid, double *box)
{
    if (!h || !box || (mid != 1 && mid != 2)) return LSMP_E_INVALID;
    if (!h->mk_set[mid - 1]) return LSMP_E_STATE;
    return download_field(h, h->mk[mid - 1], box);
}

int lsmp_marker_load(lsmp_handle h, int mid)
{
    if (!h || (mid != 1 && mid != 2)) return LSMP_E_INVALID;
    if (!h->mk_set[mid - 1]) return LSMP_E_STATE;
    LSMP_HIP(hipMemcpyAsync(h->phi, h->mk[mid - 1], h->ncell * h->esz, hipMemcpyDeviceToDevice, h->st));
    return LSMP_OK;
}

int lsmp_marker_store(lsmp_handle h, int mid)
{
    if (!h || (mid != 1 && mid != 2)) return LSMP_E_INVALID;
    if (!h->mk[0]) return LSMP_E_STATE;
    LSMP_HIP(hipMemcpyAsync(h->mk[mid - 1], h->phi, h->ncell * h->esz, hipMemcpyDeviceToDevice, h->st));
    h->mk_set[mid - 1] = true;
    return LSMP_OK;
}

int lsmp_combine_markers(lsmp_handle h, int *reconstruct)
{
    if (!h || !reconstruct) return LSMP_E_INVALID;
    if (!h->mk_set[0] || !h->mk_set[1]) return LSMP_E_STATE;
    LSMP_HIP(hipMemsetAsync(h->mflag, 0, sizeof(unsigned), h->st));
    if (h->fp32)
        launch_combine_markers<float>(h->st, (const float *)h->mk[0], (const float *)h->mk[1], (float *)h->phi,
                                      h->ncell, h->mflag);
    else
        launch_combine_markers<double>(h->st, (const double *)h->mk[0], (const double *)h->mk[1], (double *)h->phi,
                                       h->ncell, h->mflag);
    LSMP_HIP(hipGetLastError());
    LSMP_HIP(hipMemcpyAsync(h->hbuf, h->mflag, sizeof(unsigned), hipMemcpyDeviceToHost, h->st));
    LSMP_HIP(hipStreamSynchronize(h->st));
    unsigned f;
    memcpy(&f, h->hbuf, sizeof f);
    *reconstruct = f != 0u;
    return LSMP_OK;
}

int lsmp_set_velocity(lsmp_handle h, const double *u_box, const double *v_box, const double *w_box)
{
    if (!h || !u_box || !v_box || (h->dim == 3 && !w_box)) return LSMP_E_INVALID;
    int rc = upload_field(h, h->u, u_box);
    if (!rc) rc = upload_field(h, h->v, v_box);
    if (!rc && h->dim == 3) rc = upload_field(h, h->w, w_box);
    if (!rc) h->vel_set = true;
    return rc;
}

int lsmp_set_velocity_analytic(lsmp_handle h, int field, double t, double t2s)
{
    if (!h || field < 0 || field > 2 || !(t2s > 0)) return LSMP_E_INVALID;
    if (field == LSMP_VEL_LEVEQUE3D && h->dim != 3) return LSMP_E_INVALID;
    int rc = h->fp32 ? velocity_analytic<float>(h, field, t, t2s) : velocity_analytic<double>(h, field, t, t2s);
    if (!rc) h->vel_set = true;
    return rc;
}

int lsmp_rk3_advect(lsmp_handle h, double dt)
{
    if (!h || !(dt > 0)) return LSMP_E_INVALID;
    if (!h->vel_set) return LSMP_E_STATE;
    Timer tm(h, 0, nullptr, true);
    int rc = h->fp32 ? advect<float>(h, dt) : advect<double>(h, dt);
    tm.stop();
    h->stats.n_advect++;
    resolve(h, false);
    return rc;
}

int lsmp_reinit(lsmp_handle h, int btn, double timestop, double *red_error, int *n_pseudo)
{
    if (!h || (btn != 0 && btn != 1)) return LSMP_E_INVALID;
    Timer tm(h, 1, nullptr, true);
    int rc = lsmp_reinit_run(h, btn, timestop, 0, true, red_error, n_pseudo);
    tm.stop();
    h->stats.n_reinit++;
    resolve(h, false);
    return rc;
}

int lsmp_mass_vol(lsmp_handle h, double *mass, double *vol)
{
    if (!h) return LSMP_E_INVALID;
    return h->fp32 ? mass_vol<float>(h, mass, vol) : mass_vol<double>(h, mass, vol);
}

int lsmp_mass_fix(lsmp_handle h, double imass, double *mass, int *n_iter)
{
    if (!h || !(imass > 0)) return LSMP_E_INVALID;
    Timer tm(h, 2, nullptr, true);
    int rc = h->fp32 ? mass_fix<float>(h, imass, mass, n_iter) : mass_fix<double>(h, imass, mass, n_iter);
    tm.stop();
    h->stats.n_massfix++;
    resolve(h, false);
    return rc;
}

int lsmp_energy_momentum(lsmp_handle h, double mu_12, int vof_marker, double *out)
{
    if (!h || !out) return LSMP_E_INVALID;
    return h->fp32 ? energy_momentum<float>(h, mu_12, vof_marker != 0, out)
                   : energy_momentum<double>(h, mu_12, vof_marker != 0, out);
}

int lsmp_interface_l1(lsmp_handle h, double *l1)
{
    if (!h || !l1) return LSMP_E_INVALID;
    if (!h->phi0_set) return LSMP_E_STATE;
    return h->fp32 ? interface_l1<float>(h, l1) : interface_l1<double>(h, l1);
}

int lsmp_get_stats(lsmp_handle h, lsmp_stats *out)
{
    if (!h || !out) return LSMP_E_INVALID;
    LSMP_HIP(hipStreamSynchronize(h->st));
    resolve(h, true);
    *out = h->stats;
    return LSMP_OK;
}

int lsmp_set_profiling(lsmp_handle h, int on)
{
    if (!h) return LSMP_E_INVALID;
    LSMP_HIP(hipStreamSynchronize(h->st));
    resolve(h, true);
    h->profiling = on != 0;
    h->ktime.clear();
    return LSMP_OK;
}

int lsmp_kernel_time(lsmp_handle h, const char *name, long long *launches, double *ms)
{
    if (!h || !name) return LSMP_E_INVALID;
    LSMP_HIP(hipStreamSynchronize(h->st));
    resolve(h, true);
    auto it = h->ktime.find(name);
    if (launches) *launches = it == h->ktime.end() ? 0 : it->second.first;
    if (ms) *ms = it == h->ktime.end() ? 0.0 : it->second.second;
    return LSMP_OK;
}

int lsmp_synchronize(lsmp_handle h)
{
    if (!h) return LSMP_E_INVALID;
    LSMP_HIP(hipStreamSynchronize(h->st));
    return LSMP_OK;
}

int lsmp_set_band_lines(lsmp_handle h, int mode)
{
    if (!h || mode < 0 || mode > 2) return LSMP_E_INVALID;
    h->band_lines = mode;
    return LSMP_OK;
}

int lsmp_set_narrow_band(lsmp_handle h, double width_cells)
{
    if (!h || !(width_cells >= 0.0) || (width_cells > 0.0 && (h->dim != 3 || width_cells < 4.0)))
        return LSMP_E_INVALID;
    h->nb_cells = width_cells;
    return LSMP_OK;
}

int lsmp_band_line_counts(lsmp_handle h, int counts[3])
{
    if (!h || !counts) return LSMP_E_INVALID;
    for (int d = 0; d < 3; ++d) counts[d] = h->last_lines[d];
    return LSMP_OK;
}

int lsmp_set_advection_scheme(lsmp_handle h, int scheme)
{
    if (!h || (scheme != LSMP_SCHEME_CRWENO && scheme != LSMP_SCHEME_UCCD)) return LSMP_E_INVALID;
    if (scheme == LSMP_SCHEME_UCCD && !h->ucoef[0]) {
        const int ns[3] = {h->nx, h->ny, h->nz};
        for (int d = 0; d < h->dim; ++d) {
            const int nb = ns[d] + 2 * GHC;
            std::vector<double> all;
            for (int kind = 1; kind <= 2; ++kind) {
                std::vector<double> tab;
                ccd_tables(nb, h->p.dx, tab, kind);   // every UCCD solver uses dx (ccd_solvers.f90:406-408)
                all.insert(all.end(), tab.begin(), tab.end());
            }
            LSMP_HIP(hipMalloc(&h->ucoef[d], all.size() * h->esz));
            if (h->fp32) {
                std::vector<float> tf(all.begin(), all.end());
                LSMP_HIP(hipMemcpy(h->ucoef[d], tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
            } else {
                LSMP_HIP(hipMemcpy(h->ucoef[d], all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
            }
        }
    }
    h->scheme = scheme;
    return LSMP_OK;
}

// ------------------------------------------------------------------ CLSVOF
int lsmp_upload_vof(lsmp_handle h, const double *vof_box)
{
    if (!h || !vof_box) return LSMP_E_INVALID;
    int rc = ensure_clsvof(h);
    if (!rc) rc = upload_field(h, h->vof, vof_box);
    if (!rc) h->vof_set = true;
    return rc;
}

int lsmp_download_vof(lsmp_handle h, double *vof_box)
{
    if (!h || !vof_box) return LSMP_E_INVALID;
    if (!h->vof_set) return LSMP_E_STATE;
    return download_field(h, h->vof, vof_box);
}

int lsmp_store_vof_old(lsmp_handle h)
{
    if (!h) return LSMP_E_INVALID;
    if (!h->vof_set) return LSMP_E_STATE;
    LSMP_HIP(hipMemcpyAsync(h->vof_old, h->vof, h->ncell * h->esz, hipMemcpyDeviceToDevice, h->st));
    h->vof_old_set = true;
    return LSMP_OK;
}

int lsmp_store_vof0(lsmp_handle h)
{
    if (!h) return LSMP_E_INVALID;
    if (!h->vof_set) return LSMP_E_STATE;
    LSMP_HIP(hipMemcpyAsync(h->vof0, h->vof, h->ncell * h->esz, hipMemcpyDeviceToDevice, h->st));
    LSMP_HIP(hipStreamSynchronize(h->st));
    h->vof0_set = true;
    return LSMP_OK;
}

int lsmp_vof_mass_vol(lsmp_handle h, double *mass, double *vol)
{
    if (!h || !mass || !vol) return LSMP_E_INVALID;
    if (!h->vof_set) return LSMP_E_STATE;
    return h->fp32 ? vof_sums<float>(h, 0, vol, mass) : vof_sums<double>(h, 0, vol, mass);
}

int lsmp_vof_interface_l1(lsmp_handle h, double *l1)
{
    if (!h || !l1) return LSMP_E_INVALID;
    if (!h->vof_set || !h->vof0_set) return LSMP_E_STATE;
    return h->fp32 ? vof_sums<float>(h, 1, l1, nullptr) : vof_sums<double>(h, 1, l1, nullptr);
}

int lsmp_set_face_velocity(lsmp_handle h, const double *u_box, const double *v_box, const double *w_box)
{
    if (!h || !u_box || !v_box || (h->dim == 3 && !w_box)) return LSMP_E_INVALID;
    int rc = ensure_clsvof(h);
    if (!rc) rc = upload_field(h, h->uf, u_box);
    if (!rc) rc = upload_field(h, h->vf, v_box);
    if (!rc && h->dim == 3) rc = upload_field(h, h->wf, w_box);
    if (!rc) h->face_set = true;
    return rc;
}

int lsmp_set_face_velocity_analytic(lsmp_handle h, int field, double t, double t2s)
{
    if (!h || field < 0 || field > 2 || !(t2s > 0)) return LSMP_E_INVALID;
    if (field == LSMP_VEL_LEVEQUE3D && h->dim != 3) return LSMP_E_INVALID;
    int rc = ensure_clsvof(h);
    if (!rc) rc = h->fp32 ? face_velocity_analytic<float>(h, field, t, t2s) : face_velocity_analytic<double>(h, field, t, t2s);
    if (!rc) h->face_set = true;
    return rc;
}

int lsmp_vof_wlic(lsmp_handle h, double dt)
{
    if (!h || !(dt > 0)) return LSMP_E_INVALID;
    if (!h->vof_set || !h->vof_old_set || !h->face_set || h->nb_cells > 0.0) return LSMP_E_STATE;
    Timer tm(h, 0, nullptr, true);
    int rc = h->fp32 ? vof_wlic<float>(h, dt) : vof_wlic<double>(h, dt);
    tm.stop();
    resolve(h, false);
    return rc;
}

int lsmp_clsvof_recon(lsmp_handle h, int *n_pseudo0, int *n_pseudo1)
{
    if (!h) return LSMP_E_INVALID;
    if (!h->vof_set || h->nb_cells > 0.0) return LSMP_E_STATE;
    Timer tm(h, 1, nullptr, true);
    int a = 0, b = 0;
    int rc = h->fp32 ? clsvof_recon<float>(h, &a, &b) : clsvof_recon<double>(h, &a, &b);
    tm.stop();
    h->stats.n_reinit++;
    resolve(h, false);
    if (n_pseudo0) *n_pseudo0 = a;
    if (n_pseudo1) *n_pseudo1 = b;
    return rc;
}

}  // extern "C"
