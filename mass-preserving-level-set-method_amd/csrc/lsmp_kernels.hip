// lsmp_kernels.hip -- gfx950 kernels of the level-set transport step.
//
// Device layout: every field holds interior cells only, index
// ((k*ny + j)*nx + i), 0-based, i fastest.  The reference's zero-gradient
// ghost fill (Disk/lib/data/branches.f90:283-310: x faces, then y, then z)
// leaves ghost (i,j,k) equal to the interior cell (clamp i, clamp j, clamp k),
// so kernels read ghosts by clamping indices and no ghost layer is stored.
#include "lsmp_device.h"
#include "lsmp_internal.h"

namespace lsmp {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
template <typename T> LSMP_DEV double to_d(T v) { return (double)v; }

// max of non-negative doubles through their bit patterns (monotone for >= 0)
LSMP_DEV void atomic_max_nonneg(unsigned long long *p, double v)
{
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

// ---------------------------------------------------------------------------
// Advection: one CRWENO5 line solve per lane (level_set_rk3_solver.f90:143-219,
// crweno_flux_split :347-414, wenojs_flux_split :288-345, Thomas TDMA.f90:1-36).
//
// Fortran cells c = 1..n of a line map to 0-based q = c-1; interface m
// (0..n) is the face between cells m and m+1.  The two boundary faces come
// from WENO-JS, faces 1..n-1 from the two tridiagonal systems.  The Thomas
// forward sweep stores (c', d') of both systems for rows 1..n-2 in scr at
// scr[((m-1)*4 + comp)*nlines + L]; the backward sweep emits the flux
// difference of each cell, from cell n down to cell 1.
// ---------------------------------------------------------------------------
template <typename T, class Loader, class Emit>
LSMP_DEV void crweno_line(const Loader &ldc, const Emit &emit, T *__restrict__ scr, size_t nlines,
                          size_t L, int n, T inv_unused, T dd, bool store)
{
    (void)inv_unused;
    const T inv_dd = T(1) / dd;
    T g[6], f[6];
    // boundary faces by WENO-JS
    T gA, fA, gB, fB, gC, fC, gD, fD, gE, fE, gF, fF;
    ldc(-2, gA, fA); ldc(-1, gB, fB); ldc(0, gC, fC); ldc(1, gD, fD); ldc(2, gE, fE); ldc(3, gF, fF);
    const T gm0 = wenojs_g(gA, gB, gC, gD, gE);
    const T fp0 = wenojs_f(fB, fC, fD, fE, fF);
    ldc(n - 2, gA, fA); ldc(n - 1, gB, fB); ldc(n, gC, fC); ldc(n + 1, gD, fD); ldc(n + 2, gE, fE);
    ldc(n + 3, gF, fF);
    const T gmN = wenojs_g(gA, gB, gC, gD, gE);
    const T fpN = wenojs_f(fB, fC, fD, fE, fF);

    // window: g[t], f[t] hold cell m-2+t
#pragma unroll
    for (int t = 0; t < 6; ++t) ldc(-1 + t, g[t], f[t]);
    T cgp = T(0), dgp = T(0), cfp = T(0), dfp = T(0);
    for (int m = 1; m <= n - 1; ++m) {
        T A, B, C, S;
        crweno_g_row_fast(g[0], g[1], g[2], g[3], g[4], A, B, C, S);
        if (m == 1) S = S - A * gm0;
        if (m == n - 1) S = S - C * gmN;
        T cg, dg;
        {
            const T r = T(1) / (m == 1 ? B : B - A * cgp);
            cg = C * r;
            dg = (m == 1 ? S : S - A * dgp) * r;
        }
        crweno_f_row_fast(f[1], f[2], f[3], f[4], f[5], A, B, C, S);
        if (m == 1) S = S - A * fp0;
        if (m == n - 1) S = S - C * fpN;
        T cf, df;
        {
            const T r = T(1) / (m == 1 ? B : B - A * cfp);
            cf = C * r;
            df = (m == 1 ? S : S - A * dfp) * r;
        }
        if (m < n - 1 && store) {
            size_t o = (size_t)(m - 1) * 4 * nlines + L;
            scr[o] = cg;
            scr[o + nlines] = dg;
            scr[o + 2 * nlines] = cf;
            scr[o + 3 * nlines] = df;
        }
        cgp = cg; dgp = dg; cfp = cf; dfp = df;
        if (m < n - 1) {
#pragma unroll
            for (int t = 0; t < 5; ++t) { g[t] = g[t + 1]; f[t] = f[t + 1]; }
            ldc(m + 4, g[5], f[5]);
        }
    }
    // backward sweep: X(n-1) = d'(n-1)
    T Xg = dgp, Xf = dfp;
    emit(n, (Xg - gmN) * inv_dd, (Xf - fpN) * inv_dd);
    for (int m = n - 2; m >= 1; --m) {
        T cg = T(0), dg = T(0), cf = T(0), df = T(0);
        if (store) {
            size_t o = (size_t)(m - 1) * 4 * nlines + L;
            cg = scr[o];
            dg = scr[o + nlines];
            cf = scr[o + 2 * nlines];
            df = scr[o + 3 * nlines];
        }
        T Xg1 = dg - cg * Xg;
        T Xf1 = df - cf * Xf;
        emit(m + 1, (Xg1 - Xg) * inv_dd, (Xf1 - Xf) * inv_dd);
        Xg = Xg1; Xf = Xf1;
    }
    emit(1, (gm0 - Xg) * inv_dd, (fp0 - Xf) * inv_dd);
}

// RHS accumulation and fused RK stage (level_set_rk3_solver.f90:16-88, 207-212).
// MODE 0: R = t1 + t2 (first direction); 1: R = (R + t1) + t2;
// 2: last direction, R complete -> stage update of phi in place.
template <typename T, int MODE, int STAGE>
LSMP_DEV void adv_emit(T *__restrict__ R, T *phi, T *__restrict__ L1, T *__restrict__ L2, size_t a, T t1,
                       T t2, T dt)
{
    if (MODE == 0) {
        R[a] = t1 + t2;
    } else if (MODE == 1) {
        R[a] = (R[a] + t1) + t2;
    } else {
        T r = (R[a] + t1) + t2;
        if (STAGE == 1) {
            L1[a] = r;
            phi[a] = phi[a] + r * dt;
        } else if (STAGE == 2) {
            L2[a] = r;
            T src = (T(-3) * L1[a] + r) / T(4);
            phi[a] = phi[a] + src * dt;
        } else {
            T src = (-L1[a] - L2[a] + T(8) * r) / T(12);
            phi[a] = phi[a] + src * dt;
        }
    }
}

// Lines along y or z: one lane per line, lanes on consecutive i (coalesced).
// Line L -> base (L / nx) * ostride + L % nx, element stride estride.
template <typename T, int MODE, int STAGE>
__global__ __launch_bounds__(256) void k_adv_strided(T *phi, const T *__restrict__ vel, T *R, T *L1, T *L2,
                                                      T *__restrict__ scr, int nx, int nlines, int n,
                                                      long ostride, long estride, T dd, T dt)
{
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nlines) return;
    const size_t base = (size_t)(L / nx) * (size_t)ostride + (size_t)(L % nx);
    auto ldc = [&](int c, T &g, T &f) {
        int q = clampi(c - 1, 0, n - 1);
        size_t a = base + (size_t)q * (size_t)estride;
        flux_split(phi[a], vel[a], g, f);
    };
    auto emit = [&](int c, T t1, T t2) {
        size_t a = base + (size_t)(c - 1) * (size_t)estride;
        adv_emit<T, MODE, STAGE>(R, phi, L1, L2, a, t1, t2, dt);
    };
    crweno_line<T>(ldc, emit, scr, (size_t)nlines, (size_t)L, n, T(0), dd, true);
}

// Lines along x: 64 lanes = 64 lines (consecutive j); the contiguous x data is
// staged through an LDS tile of XCH rows so that global reads and writes stay
// coalesced.  Only the first direction, so MODE 0.
template <typename T>
__global__ __launch_bounds__(64) void k_adv_x(const T *__restrict__ phi, const T *__restrict__ vel,
                                               T *__restrict__ R, T *__restrict__ scr, int nx, int ny,
                                               int nlines, T dd)
{
    __shared__ T tp[64][XCH + 1];
    __shared__ T tu[64][XCH + 1];
    __shared__ T to[64][XCH + 1];
    const int t = threadIdx.x;
    const int L0 = blockIdx.x * 64;
    const int L = L0 + t;
    const bool act = L < nlines;
    const int n = nx;
    auto base_of = [&](int l) -> size_t { return (size_t)(l / ny) * (size_t)nx * ny + (size_t)(l % ny) * nx; };
    const size_t base = act ? base_of(L) : 0;
    int cur = -1;
    auto load_chunk = [&](int ch) {
        __syncthreads();
        for (int idx = t; idx < 64 * XCH; idx += 64) {
            int b = idx / XCH, e = idx % XCH;
            int lb = L0 + b, q = ch * XCH + e;
            if (lb < nlines && q < n) {
                size_t a = base_of(lb) + q;
                tp[b][e] = phi[a];
                tu[b][e] = vel[a];
            }
        }
        __syncthreads();
    };
    auto ldg = [&](int c, T &g, T &f) {   // direct load, boundary stencils only
        int q = clampi(c - 1, 0, n - 1);
        if (act) flux_split(phi[base + q], vel[base + q], g, f);
        else { g = T(0); f = T(0); }
    };
    auto ldc = [&](int c, T &g, T &f) {
        int q = clampi(c - 1, 0, n - 1);
        int ch = q / XCH;
        if (ch != cur) { load_chunk(ch); cur = ch; }
        flux_split(tp[t][q % XCH], tu[t][q % XCH], g, f);
    };
    int och = (n - 1) / XCH;
    auto flush = [&](int ch) {
        __syncthreads();
        for (int idx = t; idx < 64 * XCH; idx += 64) {
            int b = idx / XCH, e = idx % XCH;
            int lb = L0 + b, q = ch * XCH + e;
            if (lb < nlines && q < n) R[base_of(lb) + q] = to[b][e];
        }
        __syncthreads();
    };
    auto emit = [&](int c, T t1, T t2) {
        int q = c - 1;
        int ch = q / XCH;
        if (ch != och) { flush(och); och = ch; }
        to[t][q % XCH] = t1 + t2;
    };

    const T inv_dd = T(1) / dd;
    // boundary faces straight from global memory, the interior from the tile
    T gA, fA, gB, fB, gC, fC, gD, fD, gE, fE, gF, fF;
    ldg(-2, gA, fA); ldg(-1, gB, fB); ldg(0, gC, fC); ldg(1, gD, fD); ldg(2, gE, fE); ldg(3, gF, fF);
    const T gm0 = wenojs_g(gA, gB, gC, gD, gE);
    const T fp0 = wenojs_f(fB, fC, fD, fE, fF);
    ldg(n - 2, gA, fA); ldg(n - 1, gB, fB); ldg(n, gC, fC); ldg(n + 1, gD, fD); ldg(n + 2, gE, fE);
    ldg(n + 3, gF, fF);
    const T gmN = wenojs_g(gA, gB, gC, gD, gE);
    const T fpN = wenojs_f(fB, fC, fD, fE, fF);

    T g[6], f[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) ldc(-1 + q, g[q], f[q]);
    T cgp = T(0), dgp = T(0), cfp = T(0), dfp = T(0);
    const size_t nl = (size_t)nlines;
    for (int m = 1; m <= n - 1; ++m) {
        T A, B, C, S;
        crweno_g_row_fast(g[0], g[1], g[2], g[3], g[4], A, B, C, S);
        if (m == 1) S = S - A * gm0;
        if (m == n - 1) S = S - C * gmN;
        T cg, dg;
        {
            const T r = T(1) / (m == 1 ? B : B - A * cgp);
            cg = C * r;
            dg = (m == 1 ? S : S - A * dgp) * r;
        }
        crweno_f_row_fast(f[1], f[2], f[3], f[4], f[5], A, B, C, S);
        if (m == 1) S = S - A * fp0;
        if (m == n - 1) S = S - C * fpN;
        T cf, df;
        {
            const T r = T(1) / (m == 1 ? B : B - A * cfp);
            cf = C * r;
            df = (m == 1 ? S : S - A * dfp) * r;
        }
        if (m < n - 1 && act) {
            size_t o = (size_t)(m - 1) * 4 * nl + L;
            scr[o] = cg; scr[o + nl] = dg; scr[o + 2 * nl] = cf; scr[o + 3 * nl] = df;
        }
        cgp = cg; dgp = dg; cfp = cf; dfp = df;
        if (m < n - 1) {
#pragma unroll
            for (int q = 0; q < 5; ++q) { g[q] = g[q + 1]; f[q] = f[q + 1]; }
            ldc(m + 4, g[5], f[5]);
        }
    }
    T Xg = dgp, Xf = dfp;
    emit(n, (Xg - gmN) * inv_dd, (Xf - fpN) * inv_dd);
    for (int m = n - 2; m >= 1; --m) {
        T cg = T(0), dg = T(0), cf = T(0), df = T(0);
        if (act) {
            size_t o = (size_t)(m - 1) * 4 * nl + L;
            cg = scr[o]; dg = scr[o + nl]; cf = scr[o + 2 * nl]; df = scr[o + 3 * nl];
        }
        T Xg1 = dg - cg * Xg;
        T Xf1 = df - cf * Xf;
        emit(m + 1, (Xg1 - Xg) * inv_dd, (Xf1 - Xf) * inv_dd);
        Xg = Xg1; Xf = Xf1;
    }
    emit(1, (gm0 - Xg) * inv_dd, (fp0 - Xf) * inv_dd);
    flush(och);
}

// ---------------------------------------------------------------------------
// Reinitialisation stage, 3D (level_set_rk3_redis.f90:176-264 with the
// gradient :266-369, lambda :371-454 and ls_funs ns_solver/others.f90:1-45).
//
// A workgroup owns an interior tile RTX x RTY in (i, j) and marches over RKC
// planes in k.  Each thread of the tile + 1 ring keeps its own column
// phi(k-3..k+3) in registers (the z stencil); the current plane, with a
// 4-cell halo for the x and y stencils, is published to LDS.  grad,
// s2 = delta S0 (|grad| - 1) and s3 = |grad| delta^2 are formed on the tile +
// 1 ring and kept for three planes in LDS, so the 3x3x3 lambda quadrature
// (centre weight 51) reads LDS only.  The next plane's global loads are issued
// before the current plane is computed, hiding their latency.
// The RK stage is written in Shu-Osher form, algebraically identical to the
// reference's increments: phi1 = phi0 - dtau l(phi0);
// phi2 = 3/4 phi0 + 1/4 (phi1 - dtau l(phi1)); phi3 = 1/3 phi0 + 2/3 (phi2 - dtau l(phi2)).
// Stage 3 also reduces max |phi3 - phi0| (:46-62).
// ---------------------------------------------------------------------------
constexpr int RRING = RPX * RPY - REX * REY;   // halo-only cells of the LDS plane
static_assert(RRING <= RNT, "one thread per halo cell");

// LDS position (row, col) of halo cell number t: three rows on top, three at
// the bottom, three columns left and right of the tile + 1 ring.
LSMP_DEV void ring_pos(int t, int &r, int &c)
{
    if (t < 3 * RPX) { r = t / RPX; c = t % RPX; }
    else if (t < 6 * RPX) { r = RPY - 3 + (t - 3 * RPX) / RPX; c = (t - 3 * RPX) % RPX; }
    else if (t < 6 * RPX + 3 * REY) { const int u = t - 6 * RPX; r = 3 + u / 3; c = u % 3; }
    else { const int u = t - 6 * RPX - 3 * REY; r = 3 + u / 3; c = RPX - 3 + u % 3; }
}

template <typename T, int STAGE, int BTN>
__global__ __launch_bounds__(RNT) void k_redis_stage3d(const T *__restrict__ phi, T *__restrict__ phi_out,
                                                       const T *__restrict__ phi0, const T *__restrict__ S0,
                                                       RedisArgs ra, unsigned long long *errmax)
{
    __shared__ T sphi[RPY][RPX];
    __shared__ T ss2[REY][REX];
    __shared__ T ss3[REY][REX];
    const int nx = ra.nx, ny = ra.ny, nz = ra.nz;
    const T inv12dx = (T)ra.inv12dx, inv12dy = (T)ra.inv12dy, inv12dz = (T)ra.inv12dz;
    const T invdx = (T)ra.invdx, invdy = (T)ra.invdy, invdz = (T)ra.invdz;
    const T inv_w = (T)ra.inv_ls_wid, rdt = (T)ra.rdt;
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * RTX, j0 = blockIdx.y * RTY, k0 = blockIdx.z * RKC;
    const int k1 = min(k0 + RKC, nz);
    const bool ext = t < REX * REY;
    const int ex = t % REX, ey = t / REX;
    const int gi = clampi(i0 - 1 + ex, 0, nx - 1), gj = clampi(j0 - 1 + ey, 0, ny - 1);
    const bool inner = ext && ex >= 1 && ex <= RTX && ey >= 1 && ey <= RTY && (i0 - 1 + ex) < nx &&
                       (j0 - 1 + ey) < ny;
    // stencil centre in LDS: the clamped cell, so out-of-domain ring cells
    // take the values of the boundary cell (zero-gradient ghosts)
    const int pc = gi - (i0 - 4), pr = gj - (j0 - 4);
    const bool ring = t < RRING;
    int rr = 0, rc = 0;
    if (ring) ring_pos(t, rr, rc);
    const size_t plane = (size_t)nx * ny;
    const size_t col = (size_t)gj * nx + gi;
    const size_t rcol = (size_t)clampi(j0 - 4 + rr, 0, ny - 1) * nx + clampi(i0 - 4 + rc, 0, nx - 1);
    auto P = [&](int p) -> size_t { return (size_t)clampi(p, 0, nz - 1) * plane; };

    T w[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) w[d] = ext ? phi[P(k0 - 4 + d) + col] : T(0);
    T ring_cur = ring ? phi[P(k0 - 1) + rcol] : T(0);
    T s0_cur = ext ? S0[P(k0 - 1) + col] : T(0);
    T p0_cur = T(0);
    T grad_q = T(0), del_q = T(0), s0_q = T(0), phi_q = T(0), s2_q = T(0), s3_q = T(0);
    // in-plane 3x3 sums of s2 and s3 around this cell for planes p-2, p-1, p
    T b2m = T(0), b2c = T(0), b2p = T(0), b3m = T(0), b3c = T(0), b3p = T(0);
    double lerr = 0.0;
    for (int p = k0 - 1; p <= k1; ++p) {
        // next iteration's inputs, in flight while this plane is computed
        const T w_next = ext ? phi[P(p + 4) + col] : T(0);
        const T ring_next = ring ? phi[P(p + 1) + rcol] : T(0);
        const T s0_next = ext ? S0[P(p + 1) + col] : T(0);
        const T p0_next = (STAGE > 1 && inner && p >= k0 && p < k1) ? phi0[(size_t)p * plane + col] : T(0);

        if (ext) sphi[ey + 3][ex + 3] = w[3];
        if (ring) sphi[rr][rc] = ring_cur;
        __syncthreads();
        const bool live = p >= 0 && p < nz;
        T grad_p = T(0), del_p = T(0), s2_p = T(0), s3_p = T(0);
        const T s0_p = s0_cur, phi_p = w[3];
        if (ext && live) {
            T xs1, xs2, ys1, ys2, zs1, zs2;
            weno_hj_pm_fast(sphi[pr][pc - 3], sphi[pr][pc - 2], sphi[pr][pc - 1], w[3], sphi[pr][pc + 1],
                            sphi[pr][pc + 2], sphi[pr][pc + 3], inv12dx, invdx, xs1, xs2);
            weno_hj_pm_fast(sphi[pr - 3][pc], sphi[pr - 2][pc], sphi[pr - 1][pc], w[3], sphi[pr + 1][pc],
                            sphi[pr + 2][pc], sphi[pr + 3][pc], inv12dy, invdy, ys1, ys2);
            weno_hj_pm_fast(w[0], w[1], w[2], w[3], w[4], w[5], w[6], inv12dz, invdz, zs1, zs2);
            grad_p = godunov<T, 3>(xs1, xs2, ys1, ys2, zs1, zs2, s0_p);
            if (BTN == 1) {
                T H;
                ls_funs_fast(w[3], inv_w, H, del_p);
                s2_p = del_p * s0_p * (grad_p - T(1));
                s3_p = grad_p * del_p * del_p;
                ss2[ey][ex] = s2_p;
                ss3[ey][ex] = s3_p;
            }
        }
        __syncthreads();
        if (BTN == 1 && inner && live) {
            // 3x3 in-plane part of the lambda quadrature; the z direction is
            // summed from the planes kept in registers
            T a = T(0), b = T(0);
#pragma unroll
            for (int jj = -1; jj <= 1; ++jj)
#pragma unroll
                for (int ii = -1; ii <= 1; ++ii) {
                    a = a + ss2[ey + jj][ex + ii];
                    b = b + ss3[ey + jj][ex + ii];
                }
            b2m = b2c; b2c = b2p; b2p = a;
            b3m = b3c; b3c = b3p; b3p = b;
        } else if (p >= nz) {
            // plane nz is the ghost copy of plane nz-1
            b2m = b2c; b2c = b2p;
            b3m = b3c; b3c = b3p;
        }
        const int q = p - 1;
        if (inner && q >= k0 && q < k1) {
            T lam = T(0);
            if (BTN == 1) {
                const T lo2 = q == 0 ? b2c : b2m, lo3 = q == 0 ? b3c : b3m;   // plane -1 = plane 0
                const T a = T(51) * s2_q + ((lo2 + b2c) + b2p);
                const T b = T(51) * s3_q + ((lo3 + b3c) + b3p);
                if (fabs(b) > T(1.0e-12)) lam = a / b * del_q;
            }
            const T l = (s0_q - lam) * grad_q - s0_q;
            const size_t gidx = (size_t)q * plane + col;
            T pn;
            if (STAGE == 1) {
                pn = phi_q - rdt * l;
            } else if (STAGE == 2) {
                pn = T(0.75) * p0_cur + T(0.25) * (phi_q - rdt * l);
            } else {
                pn = p0_cur * T(1.0 / 3.0) + T(2.0 / 3.0) * (phi_q - rdt * l);
                lerr = fmax(lerr, (double)fabs(pn - p0_cur));
            }
            phi_out[gidx] = pn;
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) w[d] = w[d + 1];
        w[6] = w_next;
        ring_cur = ring_next;
        s0_cur = s0_next;
        p0_cur = p0_next;
        grad_q = grad_p; del_q = del_p; s0_q = s0_p; phi_q = phi_p; s2_q = s2_p; s3_q = s3_p;
    }
    if (STAGE == 3) {
        double m = block_max(lerr);
        if (t == 0) atomic_max_nonneg(errmax, m);
    }
}

// 2D stage: one tile, no marching (Disk/ and Vortex2D/ forms: |grad| without
// z, lambda box 3x3 with centre weight 17, Disk/...redis.f90:357-361, 434-438).
template <typename T, int STAGE, int BTN>
__global__ __launch_bounds__(RNT) void k_redis_stage2d(const T *__restrict__ phi, T *__restrict__ phi_out,
                                                       const T *__restrict__ phi0, const T *__restrict__ S0,
                                                       RedisArgs ra, unsigned long long *errmax)
{
    __shared__ T sphi[RPY][RPX];
    __shared__ T ss2[REY][REX];
    __shared__ T ss3[REY][REX];
    const int nx = ra.nx, ny = ra.ny;
    const T inv12dx = (T)ra.inv12dx, inv12dy = (T)ra.inv12dy;
    const T invdx = (T)ra.invdx, invdy = (T)ra.invdy;
    const T inv_w = (T)ra.inv_ls_wid, rdt = (T)ra.rdt;
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * RTX, j0 = blockIdx.y * RTY;
    const bool ext = t < REX * REY;
    const int ex = t % REX, ey = t / REX;
    const int gi = clampi(i0 - 1 + ex, 0, nx - 1), gj = clampi(j0 - 1 + ey, 0, ny - 1);
    const bool inner = ext && ex >= 1 && ex <= RTX && ey >= 1 && ey <= RTY && (i0 - 1 + ex) < nx &&
                       (j0 - 1 + ey) < ny;
    const int pc = gi - (i0 - 4), pr = gj - (j0 - 4);
    for (int idx = t; idx < RPX * RPY; idx += RNT) {
        int c = idx % RPX, r = idx / RPX;
        int gx = clampi(i0 - 4 + c, 0, nx - 1), gy = clampi(j0 - 4 + r, 0, ny - 1);
        sphi[r][c] = phi[(size_t)gy * nx + gx];
    }
    __syncthreads();
    T grad_p = T(0), del_p = T(0), s0_p = T(0), phi_p = T(0);
    if (ext) {
        const T c0 = sphi[pr][pc];
        T xs1, xs2, ys1, ys2;
        weno_hj_pm_fast(sphi[pr][pc - 3], sphi[pr][pc - 2], sphi[pr][pc - 1], c0, sphi[pr][pc + 1],
                        sphi[pr][pc + 2], sphi[pr][pc + 3], inv12dx, invdx, xs1, xs2);
        weno_hj_pm_fast(sphi[pr - 3][pc], sphi[pr - 2][pc], sphi[pr - 1][pc], c0, sphi[pr + 1][pc],
                        sphi[pr + 2][pc], sphi[pr + 3][pc], inv12dy, invdy, ys1, ys2);
        s0_p = S0[(size_t)gj * nx + gi];
        grad_p = godunov<T, 2>(xs1, xs2, ys1, ys2, T(0), T(0), s0_p);
        phi_p = c0;
        if (BTN == 1) {
            T H;
            ls_funs_fast(c0, inv_w, H, del_p);
            ss2[ey][ex] = del_p * s0_p * (grad_p - T(1));
            ss3[ey][ex] = grad_p * del_p * del_p;
        }
    }
    __syncthreads();
    double lerr = 0.0;
    if (inner) {
        T lam = T(0);
        if (BTN == 1) {
            T a = T(17) * ss2[ey][ex];
            T b = T(17) * ss3[ey][ex];
#pragma unroll
            for (int jj = -1; jj <= 1; ++jj)
#pragma unroll
                for (int ii = -1; ii <= 1; ++ii) {
                    a = a + ss2[ey + jj][ex + ii];
                    b = b + ss3[ey + jj][ex + ii];
                }
            if (fabs(b) > T(1.0e-12)) lam = a / b * del_p;
        }
        const T l = (s0_p - lam) * grad_p - s0_p;
        const size_t gidx = (size_t)gj * nx + gi;
        T pn;
        if (STAGE == 1) {
            pn = phi_p - rdt * l;
        } else if (STAGE == 2) {
            pn = T(0.75) * phi0[gidx] + T(0.25) * (phi_p - rdt * l);
        } else {
            const T p0 = phi0[gidx];
            pn = p0 * T(1.0 / 3.0) + T(2.0 / 3.0) * (phi_p - rdt * l);
            lerr = fmax(lerr, (double)fabs(pn - p0));
        }
        phi_out[gidx] = pn;
    }
    if (STAGE == 3) {
        double m = block_max(lerr);
        if (t == 0) atomic_max_nonneg(errmax, m);
    }
}

// S0 = 2 H(phi) - 1 (level_set_redis_init, redis.f90:95-109) or S0 = phi for
// btn = 0 (level_set_redis_stable, :123-135).
template <typename T>
__global__ void k_sign_init(const T *__restrict__ phi, T *__restrict__ S0, size_t n, T ls_wid, int raw)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    T p = phi[q];
    if (raw) { S0[q] = p; return; }
    T H, D;
    ls_funs(p, ls_wid, H, D);
    S0[q] = T(2) * H - T(1);
}

// Godunov |grad phi| of every interior cell from global loads, reduced to a
// maximum (level_set_redis_stable, redis.f90:137-153).  Used once per btn = 0
// call, so it is written for clarity rather than speed.
template <typename T, int DIM>
__global__ void k_grad_max(const T *__restrict__ phi, const T *__restrict__ S0, RedisArgs ra,
                           unsigned long long *gmax)
{
    const int nx = ra.nx, ny = ra.ny, nz = ra.nz;
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (q < (size_t)nx * ny * nz) {
        int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
        auto P = [&](int a, int b, int c) {
            return phi[((size_t)clampi(c, 0, nz - 1) * ny + clampi(b, 0, ny - 1)) * nx + clampi(a, 0, nx - 1)];
        };
        T xs1, xs2, ys1, ys2, zs1 = T(0), zs2 = T(0);
        weno_hj_pm(P(i - 3, j, k), P(i - 2, j, k), P(i - 1, j, k), P(i, j, k), P(i + 1, j, k), P(i + 2, j, k),
                   P(i + 3, j, k), (T)ra.inv12dx, (T)ra.invdx, xs1, xs2);
        weno_hj_pm(P(i, j - 3, k), P(i, j - 2, k), P(i, j - 1, k), P(i, j, k), P(i, j + 1, k), P(i, j + 2, k),
                   P(i, j + 3, k), (T)ra.inv12dy, (T)ra.invdy, ys1, ys2);
        if (DIM == 3)
            weno_hj_pm(P(i, j, k - 3), P(i, j, k - 2), P(i, j, k - 1), P(i, j, k), P(i, j, k + 1), P(i, j, k + 2),
                       P(i, j, k + 3), (T)ra.inv12dz, (T)ra.invdz, zs1, zs2);
        v = (double)godunov<T, DIM>(xs1, xs2, ys1, ys2, zs1, zs2, S0[q]);
    }
    double m = block_max(v);
    if (threadIdx.x == 0) atomic_max_nonneg(gmax, m);
}

// phi /= max|grad phi| (redis.f90:155-166); the maximum is read on the device.
template <typename T>
__global__ void k_scale_by_max(T *__restrict__ phi, size_t n, const unsigned long long *gmax)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double g = __longlong_as_double((long long)*gmax);
    phi[q] = (T)((double)phi[q] / g);
}

// ---------------------------------------------------------------------------
// CCD line solves for the mass correction (surface_norms, derivatives.f90:11-50;
// asgn_src_ccd ccd_solvers.f90:291-310; TWIN_BKS matrix_solver.f90:47-78).
// A box line has nb = n + 6 points; box point b holds interior cell
// clamp(b - 3).  coef holds the factorised tables of this direction as
// A, B, AA, BB, each (3, nb) column-major, then 1/Den of each row.  The forward sweep stores (S, SS)
// of rows 0..nb-2 in scr[(b*2 + comp)*nlines + L].
// ---------------------------------------------------------------------------
template <typename T, class Loader, class Emit>
LSMP_DEV void ccd_line(const Loader &ldq, const Emit &emit, const T *__restrict__ coef, T *__restrict__ scr,
                       size_t nlines, size_t L, int n, T dx, bool store)
{
    const int nb = n + 6;
    const T *A = coef, *B = coef + 3 * nb, *AA = coef + 6 * nb, *BB = coef + 9 * nb, *IDEN = coef + 12 * nb;
    const T idx = T(1) / dx, idx2 = T(1) / (dx * dx);
    // sliding window f(b-1), f(b), f(b+1)
    T fm1, f0, fp1;
    T S, SS;
    {
        T a0 = ldq(-3), a1 = ldq(-2), a2 = ldq(-1), a3 = ldq(0);
        S = (T(-3.5) * a0 + T(4) * a1 - T(0.5) * a2) * idx;
        SS = (T(34.0 / 3.0) * a0 - T(83.0 / 4.0) * a1 + T(10) * a2 - T(7.0 / 12.0) * a3) * idx2;
        fm1 = a0; f0 = a1; fp1 = a2;
    }
    if (store) {
        scr[L] = S;
        scr[nlines + L] = SS;
    }
    for (int b = 1; b <= nb - 2; ++b) {
        T s = T(15.0 / 16.0) * (-fm1 + fp1) * idx;
        T ss = (T(3) * fm1 - T(6) * f0 + T(3) * fp1) * idx2;
        s = s - (A[3 * b] * S + B[3 * b] * SS);
        ss = ss - (AA[3 * b] * S + BB[3 * b] * SS);
        S = s; SS = ss;
        if (store) {
            size_t o = (size_t)b * 2 * nlines + L;
            scr[o] = S;
            scr[o + nlines] = SS;
        }
        fm1 = f0; f0 = fp1; fp1 = ldq(b - 3 + 2);
    }
    {
        const int b = nb - 1;
        // one-sided closure of the last row reads box points nb-1 .. nb-4
        T e0 = ldq(b - 3), e1 = ldq(b - 4), e2 = ldq(b - 5), e3 = ldq(b - 6);
        T s = -(T(-3.5) * e0 + T(4) * e1 - T(0.5) * e2) * idx;
        T ss = (T(34.0 / 3.0) * e0 - T(83.0 / 4.0) * e1 + T(10) * e2 - T(7.0 / 12.0) * e3) * idx2;
        s = s - (A[3 * b] * S + B[3 * b] * SS);
        ss = ss - (AA[3 * b] * S + BB[3 * b] * SS);
        const T iden = IDEN[b];
        T solS = -(B[3 * b + 1] * ss - BB[3 * b + 1] * s) * iden;
        T solSS = (A[3 * b + 1] * ss - AA[3 * b + 1] * s) * iden;
        S = solS; SS = solSS;
        if (b - 3 >= 0 && b - 3 < n) emit(b - 3, S);
    }
    for (int b = nb - 2; b >= 0; --b) {
        T s = T(0), ss = T(0);
        if (store) {
            size_t o = (size_t)b * 2 * nlines + L;
            s = scr[o];
            ss = scr[o + nlines];
        }
        s = s - (A[3 * b + 2] * S + B[3 * b + 2] * SS);
        ss = ss - (AA[3 * b + 2] * S + BB[3 * b + 2] * SS);
        const T iden = IDEN[b];
        T solS = -(B[3 * b + 1] * ss - BB[3 * b + 1] * s) * iden;
        T solSS = (A[3 * b + 1] * ss - AA[3 * b + 1] * s) * iden;
        S = solS; SS = solSS;
        if (b - 3 >= 0 && b - 3 < n) emit(b - 3, S);
    }
}

// Lines along y or z (lanes on consecutive i).  FINAL = true fuses the mass
// correction's reductions (mass_preserving_level_set.f90:20-44 and ls_mv
// ns_solver/others.f90:191-205): |grad| from (nx, ny[, nz]), H, delta, and
// per-block partial sums of vol, mass and the lambda denominator.

template <typename T, int DIM, bool FINAL>
__global__ __launch_bounds__(256) void k_ccd_strided(const T *__restrict__ phi, const T *__restrict__ coef,
                                                      T *__restrict__ scr, T *__restrict__ out,
                                                      const T *__restrict__ dxv, const T *__restrict__ dyv,
                                                      T *__restrict__ gr, double *__restrict__ partial, int nx,
                                                      int nlines, int n, long ostride, long estride, CcdArgs ca)
{
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = L < nlines;
    double acc[3] = {0.0, 0.0, 0.0};
    if (act) {
        const size_t base = (size_t)(L / nx) * (size_t)ostride + (size_t)(L % nx);
        auto ldq = [&](int q) -> T { return phi[base + (size_t)clampi(q, 0, n - 1) * (size_t)estride]; };
        auto emit = [&](int q, T d) {
            const size_t a = base + (size_t)q * (size_t)estride;
            if (!FINAL) {
                out[a] = d;
            } else {
                T g;
                if (DIM == 3) g = sqrt(sq(dxv[a]) + sq(dyv[a]) + sq(d));
                else g = sqrt(sq(dxv[a]) + sq(d));
                gr[a] = g;
                // the mass integral is formed in FP64 in both forms
                double h, dd;
                ls_funs<double>((double)phi[a], ca.ls_wid, h, dd);
                const double gg = (double)g;
                const double r12 = ca.rho_12;
                double rho, plam;
                acc[0] += h * ca.dv;
                if (!ca.inverse) rho = h + r12 * (1.0 - h);
                else rho = r12 * h + (1.0 - h);
                acc[1] += rho * h * ca.dv;
                plam = dd * dd * gg;
                if (ca.inverse) plam = plam * (2.0 * (r12 - 1.0) * h + 1.0) * ca.dx * ca.dy * ca.dz;
                else plam = plam * (2.0 * (1.0 - r12) * h + r12) * ca.dx * ca.dy * ca.dz;
                acc[2] += plam;
            }
        };
        ccd_line<T>(ldq, emit, coef, scr, (size_t)nlines, (size_t)L, n, (T)ca.dx, true);
    }
    if (FINAL) {
        block_sum<3>(acc);
        if (threadIdx.x == 0) {
            partial[3 * blockIdx.x + 0] = acc[0];
            partial[3 * blockIdx.x + 1] = acc[1];
            partial[3 * blockIdx.x + 2] = acc[2];
        }
    }
}

// Lines along x through an LDS tile (first direction, writes d phi / dx).
template <typename T>
__global__ __launch_bounds__(64) void k_ccd_x(const T *__restrict__ phi, const T *__restrict__ coef,
                                               T *__restrict__ scr, T *__restrict__ out, int nx, int ny,
                                               int nlines, T dx)
{
    __shared__ T tp[64][XCH + 1];
    __shared__ T to[64][XCH + 1];
    const int t = threadIdx.x;
    const int L0 = blockIdx.x * 64;
    const int L = L0 + t;
    const bool act = L < nlines;
    const int n = nx;
    auto base_of = [&](int l) -> size_t { return (size_t)(l / ny) * (size_t)nx * ny + (size_t)(l % ny) * nx; };
    int cur = -1;
    auto load_chunk = [&](int ch) {
        __syncthreads();
        for (int idx = t; idx < 64 * XCH; idx += 64) {
            int b = idx / XCH, e = idx % XCH;
            int lb = L0 + b, q = ch * XCH + e;
            if (lb < nlines && q < n) tp[b][e] = phi[base_of(lb) + q];
        }
        __syncthreads();
    };
    // ldq is called with non-decreasing clamped q in the forward sweep; the
    // closure of the last row re-reads the line end from global memory.
    const size_t base = act ? base_of(L) : 0;
    bool tail = false;
    auto ldq = [&](int q) -> T {
        q = clampi(q, 0, n - 1);
        if (tail) return act ? phi[base + q] : T(0);
        int ch = q / XCH;
        if (ch != cur) { load_chunk(ch); cur = ch; }
        return tp[t][q % XCH];
    };
    int och = (n - 1) / XCH;
    auto flush = [&](int ch) {
        __syncthreads();
        for (int idx = t; idx < 64 * XCH; idx += 64) {
            int b = idx / XCH, e = idx % XCH;
            int lb = L0 + b, q = ch * XCH + e;
            if (lb < nlines && q < n) out[base_of(lb) + q] = to[b][e];
        }
        __syncthreads();
    };
    auto emit = [&](int q, T d) {
        int ch = q / XCH;
        if (ch != och) { flush(och); och = ch; }
        to[t][q % XCH] = d;
    };
    // the forward part of ccd_line, inlined here so that the tail reload can
    // switch to global memory
    const int nb = n + 6;
    const T *A = coef, *B = coef + 3 * nb, *AA = coef + 6 * nb, *BB = coef + 9 * nb, *IDEN = coef + 12 * nb;
    const T idx = T(1) / dx, idx2 = T(1) / (dx * dx);
    const size_t nl = (size_t)nlines;
    T fm1, f0, fp1, S, SS;
    {
        T a0 = ldq(-3), a1 = ldq(-2), a2 = ldq(-1), a3 = ldq(0);
        S = (T(-3.5) * a0 + T(4) * a1 - T(0.5) * a2) * idx;
        SS = (T(34.0 / 3.0) * a0 - T(83.0 / 4.0) * a1 + T(10) * a2 - T(7.0 / 12.0) * a3) * idx2;
        fm1 = a0; f0 = a1; fp1 = a2;
    }
    if (act) { scr[L] = S; scr[nl + L] = SS; }
    for (int b = 1; b <= nb - 2; ++b) {
        T s = T(15.0 / 16.0) * (-fm1 + fp1) * idx;
        T ss = (T(3) * fm1 - T(6) * f0 + T(3) * fp1) * idx2;
        s = s - (A[3 * b] * S + B[3 * b] * SS);
        ss = ss - (AA[3 * b] * S + BB[3 * b] * SS);
        S = s; SS = ss;
        if (act) {
            size_t o = (size_t)b * 2 * nl + L;
            scr[o] = S; scr[o + nl] = SS;
        }
        fm1 = f0; f0 = fp1; fp1 = ldq(b - 1);
    }
    tail = true;
    {
        const int b = nb - 1;
        T e0 = ldq(b - 3), e1 = ldq(b - 4), e2 = ldq(b - 5), e3 = ldq(b - 6);
        T s = -(T(-3.5) * e0 + T(4) * e1 - T(0.5) * e2) * idx;
        T ss = (T(34.0 / 3.0) * e0 - T(83.0 / 4.0) * e1 + T(10) * e2 - T(7.0 / 12.0) * e3) * idx2;
        s = s - (A[3 * b] * S + B[3 * b] * SS);
        ss = ss - (AA[3 * b] * S + BB[3 * b] * SS);
        const T iden = IDEN[b];
        S = -(B[3 * b + 1] * ss - BB[3 * b + 1] * s) * iden;
        SS = (A[3 * b + 1] * ss - AA[3 * b + 1] * s) * iden;
    }
    for (int b = nb - 2; b >= 0; --b) {
        T s = T(0), ss = T(0);
        if (act) {
            size_t o = (size_t)b * 2 * nl + L;
            s = scr[o]; ss = scr[o + nl];
        }
        s = s - (A[3 * b + 2] * S + B[3 * b + 2] * SS);
        ss = ss - (AA[3 * b + 2] * S + BB[3 * b + 2] * SS);
        const T iden = IDEN[b];
        T solS = -(B[3 * b + 1] * ss - BB[3 * b + 1] * s) * iden;
        T solSS = (A[3 * b + 1] * ss - AA[3 * b + 1] * s) * iden;
        S = solS; SS = solSS;
        if (b - 3 >= 0 && b - 3 < n) emit(b - 3, S);
    }
    flush(och);
}

// Deterministic final reduction of per-block partials (fixed order), then the
// mass-correction multiplier lambda = (imass - mass) / sum (mass_preserving_level_set.f90:46).
// scal: [0] vol, [1] mass, [2] denominator, [3] lambda, [4] flag (1 = empty band).
__global__ void k_finalize3(const double *__restrict__ partial, int nblk, double imass, double *scal)
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        acc[0] += partial[3 * b + 0];
        acc[1] += partial[3 * b + 1];
        acc[2] += partial[3 * b + 2];
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        scal[0] = acc[0];
        scal[1] = acc[1];
        scal[2] = acc[2];
        if (acc[2] != 0.0) {
            scal[3] = (imass - acc[1]) / acc[2];
            scal[4] = 0.0;
        } else {
            scal[3] = 0.0;
            scal[4] = 1.0;
        }
    }
}

// phi += lambda delta |grad phi| (mass_preserving_level_set.f90:52-58); lambda
// is read from device memory so the host need not round-trip it.
template <typename T>
__global__ void k_mass_update(T *__restrict__ phi, const T *__restrict__ gr, size_t n, T ls_wid,
                              const double *__restrict__ scal)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double lam = scal[3];
    double H, D;
    ls_funs<double>((double)phi[q], (double)ls_wid, H, D);
    phi[q] = (T)((double)phi[q] + lam * D * (double)gr[q]);
}

// LS mass and volume (ls_mv, ns_solver/others.f90:185-212) as block partials.
template <typename T>
__global__ void k_mass_vol(const T *__restrict__ phi, size_t n, CcdArgs ca, double *__restrict__ partial)
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double h, D;
        ls_funs<double>((double)phi[q], ca.ls_wid, h, D);
        const double r12 = ca.rho_12;
        double rho;
        acc[0] += h * ca.dv;
        if (!ca.inverse) rho = h + r12 * (1.0 - h);
        else rho = r12 * h + (1.0 - h);
        acc[1] += rho * h * ca.dv;
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = acc[0];
        partial[3 * blockIdx.x + 1] = acc[1];
        partial[3 * blockIdx.x + 2] = 0.0;
    }
}

// Shape metric partial sums of |H(phi0/eps) - H(phi/eps)| (output.f90:178-204).
template <typename T>
__global__ void k_l1(const T *__restrict__ phi, const T *__restrict__ phi0, size_t n, double ls_wid,
                     double *__restrict__ partial)
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double src = heavyside<double>((double)phi0[q] / ls_wid);
        double tgt = heavyside<double>((double)phi[q] / ls_wid);
        acc[0] += fabs(src - tgt);
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = acc[0];
        partial[3 * blockIdx.x + 1] = 0.0;
        partial[3 * blockIdx.x + 2] = 0.0;
    }
}

// Prescribed velocity on the device (velocity.f90): cell centres
// x(i) = xstart + (i - 0.5) dx with (i - 0.5) formed in single precision
// (Vortex/lib/data/tree.f90:254-256).

template <typename T>
__global__ void k_velocity(T *__restrict__ u, T *__restrict__ v, T *__restrict__ w, VelArgs va)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)va.nx * va.ny * va.nz;
    if (q >= n) return;
    const int i = (int)(q % va.nx), j = (int)((q / va.nx) % va.ny), k = (int)(q / ((size_t)va.nx * va.ny));
    const double pi = 3.141592653589793238462643383279502884;
    const double x = va.xstart + (double)((float)(i + 1) - 0.5f) * va.dx;
    const double y = va.ystart + (double)((float)(j + 1) - 0.5f) * va.dy;
    const double z = va.zstart + (double)((float)(k + 1) - 0.5f) * va.dz;
    double uu, vv, ww = 0.0;
    if (va.field == 0) {
        uu = -va.omega * (y - 0.5);
        vv = va.omega * (x - 0.5);
    } else if (va.field == 1) {
        uu = -sq(sin(pi * x)) * sin(2.0 * pi * y) * va.ct;
        vv = sq(sin(pi * y)) * sin(2.0 * pi * x) * va.ct;
    } else {
        uu = 2.0 * sq(sin(pi * x)) * sin(2.0 * pi * y) * sin(2.0 * pi * z) * va.ct;
        vv = -sin(2.0 * pi * x) * sq(sin(pi * y)) * sin(2.0 * pi * z) * va.ct;
        ww = -sin(2.0 * pi * x) * sin(2.0 * pi * y) * sq(sin(pi * z)) * va.ct;
    }
    u[q] = (T)uu;
    v[q] = (T)vv;
    w[q] = (T)ww;
}

template <typename TO, typename TI>
__global__ void k_convert(TO *__restrict__ out, const TI *__restrict__ in, size_t n)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = (TO)in[q];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
void launch_adv_pass(hipStream_t st, int dir, int mode, int stage, T *phi, const T *vel, T *R, T *L1, T *L2,
                     T *scr, int nx, int ny, int nz, T dd, T dt)
{
    if (dir == 0) {
        const int nlines = ny * nz;
        hipLaunchKernelGGL((k_adv_x<T>), dim3(nblk(nlines, 64)), dim3(64), 0, st, phi, vel, R, scr, nx, ny, nlines,
                           dd);
        return;
    }
    const int n = dir == 1 ? ny : nz;
    const int nlines = dir == 1 ? nx * nz : nx * ny;
    const long ostride = dir == 1 ? (long)nx * ny : (long)nx;
    const long estride = dir == 1 ? (long)nx : (long)nx * ny;
    const dim3 g(nblk(nlines, 256)), b(256);
#define LSMP_ADV(M, S)                                                                                          \
    hipLaunchKernelGGL((k_adv_strided<T, M, S>), g, b, 0, st, phi, vel, R, L1, L2, scr, nx, nlines, n, ostride, \
                       estride, dd, dt)
    if (mode == 1) LSMP_ADV(1, 1);
    else if (stage == 1) LSMP_ADV(2, 1);
    else if (stage == 2) LSMP_ADV(2, 2);
    else LSMP_ADV(2, 3);
#undef LSMP_ADV
}

template <typename T>
void launch_redis_stage(hipStream_t st, int dim, int stage, int btn, const T *phi, T *phi_out, const T *phi0,
                        const T *S0, const RedisArgs &ra, unsigned long long *errmax)
{
    if (dim == 3) {
        const dim3 g(nblk(ra.nx, RTX), nblk(ra.ny, RTY), nblk(ra.nz, RKC));
#define LSMP_R3(S, B) hipLaunchKernelGGL((k_redis_stage3d<T, S, B>), g, dim3(RNT), 0, st, phi, phi_out, phi0, S0, ra, errmax)
        if (btn) { if (stage == 1) LSMP_R3(1, 1); else if (stage == 2) LSMP_R3(2, 1); else LSMP_R3(3, 1); }
        else { if (stage == 1) LSMP_R3(1, 0); else if (stage == 2) LSMP_R3(2, 0); else LSMP_R3(3, 0); }
#undef LSMP_R3
    } else {
        const dim3 g(nblk(ra.nx, RTX), nblk(ra.ny, RTY), 1);
#define LSMP_R2(S, B) hipLaunchKernelGGL((k_redis_stage2d<T, S, B>), g, dim3(RNT), 0, st, phi, phi_out, phi0, S0, ra, errmax)
        if (btn) { if (stage == 1) LSMP_R2(1, 1); else if (stage == 2) LSMP_R2(2, 1); else LSMP_R2(3, 1); }
        else { if (stage == 1) LSMP_R2(1, 0); else if (stage == 2) LSMP_R2(2, 0); else LSMP_R2(3, 0); }
#undef LSMP_R2
    }
}

template <typename T>
void launch_sign_init(hipStream_t st, const T *phi, T *S0, size_t n, T ls_wid, int raw)
{
    hipLaunchKernelGGL((k_sign_init<T>), dim3(nblk(n, 256)), dim3(256), 0, st, phi, S0, n, ls_wid, raw);
}

template <typename T>
void launch_grad_max(hipStream_t st, int dim, const T *phi, const T *S0, const RedisArgs &ra,
                     unsigned long long *gmax)
{
    const size_t n = (size_t)ra.nx * ra.ny * ra.nz;
    if (dim == 3)
        hipLaunchKernelGGL((k_grad_max<T, 3>), dim3(nblk(n, 256)), dim3(256), 0, st, phi, S0, ra, gmax);
    else
        hipLaunchKernelGGL((k_grad_max<T, 2>), dim3(nblk(n, 256)), dim3(256), 0, st, phi, S0, ra, gmax);
}

template <typename T>
void launch_scale_by_max(hipStream_t st, T *phi, size_t n, const unsigned long long *gmax)
{
    hipLaunchKernelGGL((k_scale_by_max<T>), dim3(nblk(n, 256)), dim3(256), 0, st, phi, n, gmax);
}

template <typename T>
void launch_ccd_pass(hipStream_t st, int dir, int dim, bool final_pass, const T *phi, const T *coef, T *scr, T *out,
                     const T *dxv, const T *dyv, T *gr, double *partial, int nx, int ny, int nz, const CcdArgs &ca,
                     int *nblocks_out)
{
    if (dir == 0) {
        const int nlines = ny * nz;
        hipLaunchKernelGGL((k_ccd_x<T>), dim3(nblk(nlines, 64)), dim3(64), 0, st, phi, coef, scr, out, nx, ny,
                           nlines, (T)ca.dx);
        if (nblocks_out) *nblocks_out = 0;
        return;
    }
    const int n = dir == 1 ? ny : nz;
    const int nlines = dir == 1 ? nx * nz : nx * ny;
    const long ostride = dir == 1 ? (long)nx * ny : (long)nx;
    const long estride = dir == 1 ? (long)nx : (long)nx * ny;
    const dim3 g(nblk(nlines, 256)), b(256);
    if (!final_pass)
        hipLaunchKernelGGL((k_ccd_strided<T, 3, false>), g, b, 0, st, phi, coef, scr, out, dxv, dyv, gr, partial, nx,
                           nlines, n, ostride, estride, ca);
    else if (dim == 3)
        hipLaunchKernelGGL((k_ccd_strided<T, 3, true>), g, b, 0, st, phi, coef, scr, out, dxv, dyv, gr, partial, nx,
                           nlines, n, ostride, estride, ca);
    else
        hipLaunchKernelGGL((k_ccd_strided<T, 2, true>), g, b, 0, st, phi, coef, scr, out, dxv, dyv, gr, partial, nx,
                           nlines, n, ostride, estride, ca);
    if (nblocks_out) *nblocks_out = (int)g.x;
}

void launch_finalize3(hipStream_t st, const double *partial, int nblk_, double imass, double *scal)
{
    hipLaunchKernelGGL(k_finalize3, dim3(1), dim3(1024), 0, st, partial, nblk_, imass, scal);
}

template <typename T>
void launch_mass_update(hipStream_t st, T *phi, const T *gr, size_t n, T ls_wid, const double *scal)
{
    hipLaunchKernelGGL((k_mass_update<T>), dim3(nblk(n, 256)), dim3(256), 0, st, phi, gr, n, ls_wid, scal);
}

template <typename T>
void launch_mass_vol(hipStream_t st, const T *phi, size_t n, const CcdArgs &ca, double *partial, int nblocks)
{
    hipLaunchKernelGGL((k_mass_vol<T>), dim3(nblocks), dim3(256), 0, st, phi, n, ca, partial);
}

template <typename T>
void launch_l1(hipStream_t st, const T *phi, const T *phi0, size_t n, double ls_wid, double *partial, int nblocks)
{
    hipLaunchKernelGGL((k_l1<T>), dim3(nblocks), dim3(256), 0, st, phi, phi0, n, ls_wid, partial);
}

template <typename T>
void launch_velocity(hipStream_t st, T *u, T *v, T *w, const VelArgs &va)
{
    const size_t n = (size_t)va.nx * va.ny * va.nz;
    hipLaunchKernelGGL((k_velocity<T>), dim3(nblk(n, 256)), dim3(256), 0, st, u, v, w, va);
}

template <typename TO, typename TI>
void launch_convert(hipStream_t st, TO *out, const TI *in, size_t n)
{
    hipLaunchKernelGGL((k_convert<TO, TI>), dim3(nblk(n, 256)), dim3(256), 0, st, out, in, n);
}

void launch_zero_u64(hipStream_t st, unsigned long long *p, int n)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, n);
}

#define LSMP_INST(T)                                                                                           \
    template void launch_adv_pass<T>(hipStream_t, int, int, int, T *, const T *, T *, T *, T *, T *, int, int, int, \
                                     T, T);                                                                    \
    template void launch_redis_stage<T>(hipStream_t, int, int, int, const T *, T *, const T *, const T *,          \
                                        const RedisArgs &, unsigned long long *);                              \
    template void launch_sign_init<T>(hipStream_t, const T *, T *, size_t, T, int);                            \
    template void launch_grad_max<T>(hipStream_t, int, const T *, const T *, const RedisArgs &,                \
                                     unsigned long long *);                                                    \
    template void launch_scale_by_max<T>(hipStream_t, T *, size_t, const unsigned long long *);                \
    template void launch_ccd_pass<T>(hipStream_t, int, int, bool, const T *, const T *, T *, T *, const T *,     \
                                     const T *, T *, double *, int, int, int, const CcdArgs &, int *);           \
    template void launch_mass_update<T>(hipStream_t, T *, const T *, size_t, T, const double *);               \
    template void launch_mass_vol<T>(hipStream_t, const T *, size_t, const CcdArgs &, double *, int);           \
    template void launch_l1<T>(hipStream_t, const T *, const T *, size_t, double, double *, int);              \
    template void launch_velocity<T>(hipStream_t, T *, T *, T *, const VelArgs &);

LSMP_INST(double)
LSMP_INST(float)
template void launch_convert<double, float>(hipStream_t, double *, const float *, size_t);
template void launch_convert<float, double>(hipStream_t, float *, const double *, size_t);

}  // namespace lsmp
