// lsmp_redis.hip -- the 3D reinitialisation stage (level_set_rk3_redis.f90:176-264
// with the gradient :266-369, lambda :371-454 and ls_funs ns_solver/others.f90:1-45),
// as two kernels per RK stage.
//
// k_redis3d updates every cell of the domain with lambda = 0.  That is the
// final value wherever delta = 0 (lambda is multiplied by delta), i.e. outside
// the band |phi| < ls_wid.  For the cells inside it, it also stores |grad| in
// a side field and lists the cell: first in the workgroup's list in LDS,
// which the workgroup appends to the device list with one atomic at its end.
// k_redis3d_band then takes the listed cells, four lanes per cell over the
// whole device, forms s2 = delta S0 (|grad| - 1) and s3 = |grad| delta^2 of each
// band cell of the 3x3x3 box around it, lambda from their sums (centre weight
// 51) and the cell's final value with the same expression.  Off the band s2
// and s3 are zero, so the box sums take band cells only.
//
// k_redis3d: a workgroup of R3X*R3Y threads owns an interior tile in (i, j),
// one cell per thread, and marches over KC planes in k.  Each thread keeps
// its column phi(k-3..k+3) in registers; the current plane, with a 4-cell
// halo for the x and y stencils, goes to LDS.  The WENO-HJ correction terms
// of D+ at m and of D- at m+1 are built from the same four second
// differences s(m-1..m+2) and share their smoothness indicators, their
// squares and the pairwise products of those (in reversed roles), so each
// direction evaluates one face m+1/2 per cell.  z: each plane evaluates face
// p+1/2 and carries its D- term and the second differences to the next
// plane.  x and y: each thread evaluates the faces i+1/2 and j+1/2 of its
// cell and hands the D- terms to its neighbours through LDS; one wave (a
// different one from tile to tile) also evaluates the tile's first faces
// i0-1/2 and j0-1/2.  Two barriers per
// plane: the plane is in LDS, the D- terms are.
//
// The RK stage is written in Shu-Osher form, algebraically identical to the
// reference's increments: phi1 = phi0 - dtau l(phi0);
// phi2 = 3/4 phi0 + 1/4 (phi1 - dtau l(phi1)); phi3 = 1/3 phi0 + 2/3 (phi2 - dtau l(phi2)).
// Stage 3 also reduces max |phi3 - phi0| (:46-62).
#include <hip/hip_runtime.h>

#include "lsmp_device.h"
#include "lsmp_internal.h"

namespace lsmp {

namespace {

constexpr int NT = R3X * R3Y;                   // one thread per tile cell
constexpr int NHALO = P3X * P3Y - R3X * R3Y;    // LDS plane cells outside the tile
static_assert(NHALO <= 2 * NT, "at most two halo cells per thread");
static_assert(NT % 64 == 0, "whole waves");

// LDS position (row, col) of halo cell number t of the (R3X + 8) x (R3Y + 8)
// plane: four rows on top, four at the bottom, four columns left and right.
LSMP_DEV void halo_cell(int t, int &r, int &c)
{
    if (t < 4 * P3X) { r = t / P3X; c = t % P3X; }
    else if (t < 8 * P3X) { r = P3Y - 4 + (t - 4 * P3X) / P3X; c = (t - 4 * P3X) % P3X; }
    else { const int u = t - 8 * P3X; r = 4 + u / 8; c = u % 8 < 4 ? u % 8 : P3X - 8 + u % 8; }
}

// delta != 0 only where ls_funs_fast takes its middle branch
template <typename T> LSMP_DEV bool in_band_x(T x)
{
    return !(x > T(1) - T(1.0e-12)) && !(x < T(-1) + T(1.0e-12));
}

// Godunov |grad phi| (level_set_rk3_redis.f90:341-360): with sigma the sign of
// S0, max(-min(D+, 0), max(D-, 0)) is max(-sigma D+, sigma D-, 0) per direction
// (and max(max(D+, 0), -min(D-, 0)) the same with sigma = -1).
template <typename T> LSMP_DEV T godunov_sig(T xp, T xm, T yp, T ym, T zp, T zm, T s0)
{
    const T sg = s0 >= T(0) ? T(1) : T(-1);
    const T gx = fmax(fmax(-(sg * xp), sg * xm), T(0));
    const T gy = fmax(fmax(-(sg * yp), sg * ym), T(0));
    const T gz = fmax(fmax(-(sg * zp), sg * zm), T(0));
    return sqrt_w(sq(gx) + sq(gy) + sq(gz));
}

// The stage's new value from the pseudo-time RHS l (Shu-Osher form).
template <typename T, int STAGE> LSMP_DEV T rk_stage(T phi_c, T p0, T l, T rdt)
{
    if (STAGE == 1) return phi_c - rdt * l;
    if (STAGE == 2) return T(0.75) * p0 + T(0.25) * (phi_c - rdt * l);
    return p0 * T(1.0 / 3.0) + T(2.0 / 3.0) * (phi_c - rdt * l);
}

template <typename T> LSMP_DEV T load_at(const T *plane_base, unsigned byte_off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(plane_base) + byte_off);
}

}  // namespace

template <typename T, int STAGE, int BTN, int KC>
__global__ __launch_bounds__(NT) void k_redis3d(const T *__restrict__ phi, T *__restrict__ phi_out,
                                                const T *__restrict__ phi0, const T *__restrict__ S0, RedisArgs ra,
                                                unsigned long long *ctl, const unsigned char *__restrict__ tiles,
                                                T *__restrict__ bgr, unsigned *__restrict__ blist,
                                                unsigned *__restrict__ bcount)
{
    __shared__ T sphi[P3Y][P3X];
    // D- terms of the x and y faces: column tx (row ty) holds the one of cell tx (ty)
    __shared__ T mxs[R3Y][R3X + 1];
    __shared__ T mys[R3Y + 1][R3X];
    static_assert(R3X + R3Y <= 64, "the first faces in one wave");
    // the workgroup's band cells, (plane - k0) * NT + thread, and their count
    __shared__ unsigned short lcell[BTN == 1 ? KC * NT : 1];
    __shared__ unsigned lcount, lbase;
    static_assert(KC * NT <= 65536, "local band-cell index in 16 bits");
    if (ctl[1]) return;   // the pseudo-time loop has stopped (k_redis_ctl)
    const int nx = ra.nx, ny = ra.ny, nz = ra.nz;
    // tiles in x-fastest order, renumbered so that each XCD takes a contiguous
    // run of them: the x and y halos of a tile are then mostly in its XCD's L2
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int tile_id = xcd_renumber((int)((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x),
                                     gx * gy * (int)gridDim.z);
    const int bx = tile_id % gx, by = (tile_id / gx) % gy, bz = tile_id / (gx * gy);
    // narrow band: a workgroup with no cell of the tube keeps its cells.  Stage
    // 1 writes the step's output buffer and stage 2 the one stage 3 reads, so
    // both copy; stage 3 writes the buffer stage 1 filled, which holds them.
    if (tiles && !tiles[tile_id]) {
        if (STAGE == 3) return;
        const int ci = bx * R3X + (int)threadIdx.x % R3X, cj = by * R3Y + (int)threadIdx.x / R3X;
        if (ci >= nx || cj >= ny) return;
        const size_t c0 = (size_t)cj * nx + ci, pl = (size_t)nx * ny;
        for (int k = bz * KC; k < min((bz + 1) * KC, nz); ++k)
            phi_out[(size_t)k * pl + c0] = phi[(size_t)k * pl + c0];
        return;
    }
    const T inv12dx = (T)ra.inv12dx, inv12dy = (T)ra.inv12dy, inv12dz = (T)ra.inv12dz;
    const T inv_w = (T)ra.inv_ls_wid, rdt = (T)ra.rdt, c3 = (T)ra.three;
    const int t = threadIdx.x;
    const int i0 = bx * R3X, j0 = by * R3Y, k0 = bz * KC;
    const int k1 = min(k0 + KC, nz);
    const int tx = t % R3X, ty = t / R3X;
    // cells past the domain edge hold the boundary cell (zero-gradient ghosts)
    const int gi = min(i0 + tx, nx - 1), gj = min(j0 + ty, ny - 1);
    const bool inner = i0 + tx < nx && j0 + ty < ny;
    const bool halo = t < NHALO, halo2 = t + NT < NHALO;
    int hr = 0, hc = 0, hr2 = 0, hc2 = 0;
    if (halo) halo_cell(t, hr, hc);
    if (halo2) halo_cell(t + NT, hr2, hc2);
    // byte offsets within a plane (a plane holds fewer than 2^28 cells); the
    // plane bases are uniform over the workgroup
    const size_t plane = (size_t)nx * ny;
    const unsigned col = ((unsigned)gj * nx + gi) * (unsigned)sizeof(T);
    const unsigned hcol = ((unsigned)clampi(j0 - 4 + hr, 0, ny - 1) * nx + clampi(i0 - 4 + hc, 0, nx - 1)) *
                          (unsigned)sizeof(T);
    const unsigned hcol2 = ((unsigned)clampi(j0 - 4 + hr2, 0, ny - 1) * nx + clampi(i0 - 4 + hc2, 0, nx - 1)) *
                           (unsigned)sizeof(T);
    auto base = [&](const T *f, int p) { return f + (size_t)clampi(p, 0, nz - 1) * plane; };

    // the faces i+1/2 and j+1/2 read the LDS plane at the thread's own tile
    // position: past the domain edge it holds the boundary cell, as the halo
    // does.  One wave also takes the tile's first faces -- its lanes
    // 0..R3Y-1 the x face i0-1/2 of row u, lanes R3Y..R3Y+R3X-1 the y face
    // j0-1/2 of column u-R3Y (u the lane) -- a different wave from tile to
    // tile, so that the extra work spreads over the SIMDs of a CU
    const int fu = t - 64 * (tile_id % (NT / 64));
    const bool first_x = fu >= 0 && fu < R3Y, first_y = fu >= R3Y && fu < R3Y + R3X;
    const int f_off = first_x ? (fu + 4) * P3X + 1 : (first_y ? P3X + 4 + (fu - R3Y) : 0);
    const int f_str = first_x ? 1 : P3X;

    T w[7];   // phi(p-3 .. p+3) of this cell's column
#pragma unroll
    for (int d = 0; d < 7; ++d) w[d] = load_at(base(phi, k0 - 3 + d), col);
    // z: second differences s(k0-1), s(k0), s(k0+1) and the D- term of the
    // first plane (its face k0 - 1/2)
    T za = w[0] - T(2) * w[1] + w[2];   // s(k0 - 2)
    T zb = w[1] - T(2) * w[2] + w[3];   // s(k0 - 1)
    T zc = w[2] - T(2) * w[3] + w[4];   // s(k0)
    T zminus;
    {
        T unused;
        const T zd = w[3] - T(2) * w[4] + w[5];   // s(k0 + 1)
        weno_hj_face(za, zb, zc, zd, c3, unused, zminus);
        za = zb; zb = zc; zc = zd;
    }
    T halo_cur = halo ? load_at(base(phi, k0), hcol) : T(0);
    T halo2_cur = halo2 ? load_at(base(phi, k0), hcol2) : T(0);
    if (BTN == 1 && t == 0) lcount = 0;   // read after the loop's first barrier
    double lerr = 0.0;
    for (int pl = k0; pl < k1; ++pl) {
        // the plane index is the same in every lane: plane addresses are scalar
        const int p = __builtin_amdgcn_readfirstlane(pl);
        // next plane's inputs, in flight while this one is computed
        const T w_next = load_at(base(phi, p + 4), col);
        const T halo_next = halo ? load_at(base(phi, p + 1), hcol) : T(0);
        const T halo2_next = halo2 ? load_at(base(phi, p + 1), hcol2) : T(0);
        // this plane's S0 and phi0, used after the two barriers
        const T s0_cur = load_at(base(S0, p), col);
        const T p0_cur = STAGE > 1 ? load_at(base(phi0, p), col) : T(0);

        // the previous plane's readers have passed its second barrier
        sphi[ty + 4][tx + 4] = w[3];
        if (halo) sphi[hr][hc] = halo_cur;
        if (halo2) sphi[hr2][hc2] = halo2_cur;

        // z face p + 1/2: its D+ term is this plane's, its D- term the next one's
        const T zs_new = w[4] - T(2) * w[5] + w[6];   // s(p + 2)
        T zplus, zminus_next;
        weno_hj_face(za, zb, zc, zs_new, c3, zplus, zminus_next);
        const T az = inv12dz * central12(w[1], w[2], w[4], w[5]);
        const T zs1 = az + inv12dz * zplus, zs2 = az - inv12dz * zminus;
        __syncthreads();
        // x face i+1/2: D+ term of this cell, D- term of cell i+1
        T xs1, ax, ys1, ay;
        {
            const T *r = &sphi[ty + 4][tx + 4];
            const T m2 = r[-2], m1 = r[-1], c = w[3], p1 = r[1], p2 = r[2], p3 = r[3];
            ax = inv12dx * central12(m2, m1, p1, p2);
            T plus, minus;
            weno_hj_face(m2 - T(2) * m1 + c, m1 - T(2) * c + p1, c - T(2) * p1 + p2, p1 - T(2) * p2 + p3, c3,
                         plus, minus);
            xs1 = ax + inv12dx * plus;
            mxs[ty][tx + 1] = minus;
        }
        // y face j+1/2
        {
            const T *r = &sphi[ty + 4][tx + 4];
            const T m2 = r[-2 * P3X], m1 = r[-P3X], c = w[3], p1 = r[P3X], p2 = r[2 * P3X], p3 = r[3 * P3X];
            ay = inv12dy * central12(m2, m1, p1, p2);
            T plus, minus;
            weno_hj_face(m2 - T(2) * m1 + c, m1 - T(2) * c + p1, c - T(2) * p1 + p2, p1 - T(2) * p2 + p3, c3,
                         plus, minus);
            ys1 = ay + inv12dy * plus;
            mys[ty + 1][tx] = minus;
        }
        if (first_x || first_y) {
            const T *f = &sphi[0][0] + f_off;
            const T v0 = f[0], v1 = f[f_str], v2 = f[2 * f_str], v3 = f[3 * f_str], v4 = f[4 * f_str],
                    v5 = f[5 * f_str];
            T plus, minus;
            weno_hj_face(v0 - T(2) * v1 + v2, v1 - T(2) * v2 + v3, v2 - T(2) * v3 + v4, v3 - T(2) * v4 + v5, c3,
                         plus, minus);
            if (first_x) mxs[fu][0] = minus;
            else mys[0][fu - R3Y] = minus;
        }
        __syncthreads();
        const T xs2 = ax - inv12dx * mxs[ty][tx];
        const T ys2 = ay - inv12dy * mys[ty][tx];
        const T grad = godunov_sig(xs1, xs2, ys1, ys2, zs1, zs2, s0_cur);
        // lambda = 0: final off the band, replaced by k_redis3d_band on it
        const T l = s0_cur * grad - s0_cur;
        const T pn = rk_stage<T, STAGE>(w[3], p0_cur, l, rdt);
        const bool band = BTN == 1 && inner && in_band_x(w[3] * inv_w);
        if (inner) {
            const size_t off = (size_t)p * plane;
            *reinterpret_cast<T *>(reinterpret_cast<char *>(phi_out + off) + col) = pn;
            if (band) {
                bgr[off + col / sizeof(T)] = grad;
            } else if (STAGE == 3) {
                lerr = fmax(lerr, (double)fabs(pn - p0_cur));
            }
        }
        if (BTN == 1) {
            // list this plane's band cells in LDS: one LDS atomic per wave
            const unsigned long long vote = __ballot(band);
            if (vote) {
                const int lane = t & 63, lead = __ffsll((long long)vote) - 1;
                unsigned at = 0;
                if (lane == lead) at = atomicAdd(&lcount, (unsigned)__popcll(vote));
                at = __shfl(at, lead, 64);
                if (band)
                    lcell[at + (unsigned)__popcll(vote & ((1ull << lane) - 1ull))] =
                        (unsigned short)((p - k0) * NT + t);
            }
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) w[d] = w[d + 1];
        w[6] = w_next;
        za = zb; zb = zc; zc = zs_new;
        zminus = zminus_next;
        halo_cur = halo_next;
        halo2_cur = halo2_next;
    }
    if (BTN == 1) {
        // the workgroup's list onto the device list: one atomic
        __syncthreads();
        const unsigned nl = lcount;
        if (nl) {
            if (t == 0) lbase = atomicAdd(bcount, nl);
            __syncthreads();
            const unsigned at = lbase;
            for (unsigned e = t; e < nl; e += NT) {
                const unsigned lc = lcell[e], lt = lc % NT;
                blist[at + e] = (unsigned)((size_t)(k0 + (int)(lc / NT)) * plane) +
                                (unsigned)(j0 + (int)(lt / R3X)) * (unsigned)nx + (unsigned)(i0 + (int)(lt % R3X));
            }
        }
    }
    if (STAGE == 3) {
        double m = block_max(lerr);
        if (t == 0 && m > 0.0) atomic_max_nonneg(ctl_slot(ctl, (unsigned)tile_id), m);
    }
}

// delta of ls_funs for a cell of the band (its middle branch, |x| < 1 with
// x = phi / ls_wid): 0.5 (1 + cos(pi x)) / ls_wid = cos(theta)^2 / ls_wid with
// theta = pi x / 2 in (-pi/2, pi/2), cos from its Taylor series in theta^2 to
// the term in theta^20 (remainder below 2e-17 there).  Within 2.3e-16 of
// 0.5 (1 + cos(pi x)) evaluated with a correctly rounded cosine.
template <typename T> LSMP_DEV T band_delta(T phi, T inv_w)
{
    const T th = (phi * inv_w) * T(1.570796326794896619231321691639751442);
    const T t = th * th;
    T c = T(1.0 / 2432902008176640000.0);     // 1/20!
    c = fma(c, t, T(-1.0 / 6402373705728000.0));   // -1/18!
    c = fma(c, t, T(1.0 / 20922789888000.0));      // 1/16!
    c = fma(c, t, T(-1.0 / 87178291200.0));        // -1/14!
    c = fma(c, t, T(1.0 / 479001600.0));           // 1/12!
    c = fma(c, t, T(-1.0 / 3628800.0));            // -1/10!
    c = fma(c, t, T(1.0 / 40320.0));               // 1/8!
    c = fma(c, t, T(-1.0 / 720.0));                // -1/6!
    c = fma(c, t, T(1.0 / 24.0));                  // 1/4!
    c = fma(c, t, T(-0.5));                        // -1/2!
    c = fma(c, t, T(1));
    return c * c * inv_w;
}

// The listed band cells: lambda from the 3x3x3 box sums of s2, s3 (centre
// weight 51, :409-448) in the reference's order -- in-plane sums row by row,
// then (plane k-1 + plane k) + plane k+1 -- and the cell's value.  Ghost cells
// of the box are the clamped interior cells (the zero-gradient fill of s2,
// s3).  Four lanes per listed cell: lanes 0..2 sum box planes k-1..k+1, lane
// 3 loads the cell itself and finishes it.  Consecutive stages alternate
// between two list lengths: this stage's (bcount) and the next one's (bnext),
// which this pass clears -- the previous stage's pass has read it, and the
// next stage's listing starts after this pass ends.  The stop flag, the list
// length, the list entries and the box loads of the first pass are in flight
// together: an entry at or past the length (stale, or never written) is
// clamped to a cell of the domain and its result dropped.
template <typename T, int STAGE>
__global__ __launch_bounds__(NT) void k_redis3d_band(const T *__restrict__ phi, T *__restrict__ phi_out,
                                                     const T *__restrict__ phi0, const T *__restrict__ S0,
                                                     RedisArgs ra, unsigned long long *ctl,
                                                     const T *__restrict__ bgr, const unsigned *__restrict__ blist,
                                                     const unsigned *bcount, unsigned *bnext)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *bnext = 0;
    const unsigned ncell = bcount[0];
    // stage 3: the maximum change in this workgroup's stripe, the stage
    // kernel's part of it complete now; a workgroup whose own maximum is not
    // above it skips the atomic (the 1,024 workgroups end together)
    unsigned long long *slot = ctl_slot(ctl, blockIdx.x);
    const unsigned long long stage_max = STAGE == 3 ? *slot : 0ull;
    const int nx = ra.nx, ny = ra.ny, nz = ra.nz;
    const T inv_w = (T)ra.inv_ls_wid, rdt = (T)ra.rdt;
    const unsigned plane = (unsigned)nx * (unsigned)ny, last = plane * (unsigned)nz - 1u;
    const int q = threadIdx.x & 3, g0 = (threadIdx.x & 63) & ~3;   // role, first lane of the four
    constexpr unsigned CPB = NT / 4;                                 // cells per workgroup and pass
    double lerr = 0.0;
    // the pass count is the same for every lane of the workgroup (shuffles)
    for (unsigned e0 = blockIdx.x * CPB;; e0 += gridDim.x * CPB) {
        const unsigned e = e0 + threadIdx.x / 4;
        // the list has room for every cell of the domain
        const unsigned c = min(blist[min(e, last)], last);
        // the stop flag (the stage kernel listed nothing) as a vector load,
        // behind the list entry, so that neither waits for the other
        const bool stopped = *reinterpret_cast<const volatile unsigned long long *>(ctl + 1) != 0;
        const unsigned rem = c % plane;
        const int k = (int)(c / plane), cj = (int)(rem / (unsigned)nx), ci = (int)(rem % (unsigned)nx);
        T part2 = T(0), part3 = T(0);
        T phc = T(0), s0 = T(0), grad = T(0), p0 = T(0);
        if (q < 3) {
            // box plane k-1+q: phi (band membership, delta), S0 and |grad| of
            // its nine cells, all loads in flight together; cells off the band
            // contribute nothing (their s2, s3 are zero up to sign), whatever
            // |grad| holds there
            const unsigned pb = (unsigned)clampi(k - 1 + q, 0, nz - 1) * plane;
            T pv[9], sv[9], gv[9];
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) {
                    const unsigned at = pb + (unsigned)clampi(cj - 1 + jj, 0, ny - 1) * (unsigned)nx +
                                        (unsigned)clampi(ci - 1 + ii, 0, nx - 1);
                    pv[jj * 3 + ii] = phi[at];
                    sv[jj * 3 + ii] = S0[at];
                    gv[jj * 3 + ii] = bgr[at];
                }
            if (stopped) return;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const bool inb = in_band_x(pv[r] * inv_w);
                const T dn = band_delta(pv[r], inv_w);
                const T s2 = dn * sv[r] * (gv[r] - T(1)), s3 = gv[r] * dn * dn;
                part2 = part2 + (inb ? s2 : T(0));
                part3 = part3 + (inb ? s3 : T(0));
            }
        } else {
            phc = phi[c];
            s0 = S0[c];
            grad = bgr[c];
            if (STAGE > 1) p0 = phi0[c];
            if (stopped) return;
        }
        const T a0 = __shfl(part2, g0, 64), a1 = __shfl(part2, g0 + 1, 64), a2 = __shfl(part2, g0 + 2, 64);
        const T b0 = __shfl(part3, g0, 64), b1 = __shfl(part3, g0 + 1, 64), b2 = __shfl(part3, g0 + 2, 64);
        if (q == 3 && e < ncell) {
            const T del = band_delta(phc, inv_w);
            const T c2 = del * s0 * (grad - T(1)), c3 = grad * del * del;
            const T la = T(51) * c2 + ((a0 + a1) + a2);
            const T lb = T(51) * c3 + ((b0 + b1) + b2);
            T lam = T(0);
            if (fabs(lb) > T(1.0e-12)) lam = la / lb * del;
            const T l = (s0 - lam) * grad - s0;
            const T pn = rk_stage<T, STAGE>(phc, p0, l, rdt);
            phi_out[c] = pn;
            if (STAGE == 3) lerr = fmax(lerr, (double)fabs(pn - p0));
        }
        if (e0 + gridDim.x * CPB >= ncell) break;
    }
    if (STAGE == 3) {
        double m = block_max(lerr);
        if (threadIdx.x == 0 && m > 0.0 && (unsigned long long)__double_as_longlong(m) > stage_max)
            atomic_max_nonneg(slot, m);
    }
}

static inline unsigned nb_of(int n, int b) { return (unsigned)((n + b - 1) / b); }

// workgroups of the band pass: four per CU of the device, four lanes a band
// cell (65536 cells a pass; a longer list takes more passes)
constexpr unsigned BAND_GRID = 1024;

template <typename T>
void launch_redis3d(hipStream_t st, int stage, int btn, const T *phi, T *phi_out, const T *phi0, const T *S0,
                    const RedisArgs &ra, unsigned long long *ctl, const unsigned char *tiles, const RedisBand<T> &bd,
                    bool band)
{
    // the narrow band (tiles) marches RKC_NB planes per workgroup
    const int kc = tiles ? RKC_NB : RKC;
    const dim3 g(nb_of(ra.nx, R3X), nb_of(ra.ny, R3Y), nb_of(ra.nz, kc));
#define LSMP_R3(S, B, KCV)                                                                                     \
    do {                                                                                                       \
        if (!band)                                                                                             \
            hipLaunchKernelGGL((k_redis3d<T, S, B, KCV>), g, dim3(NT), 0, st, phi, phi_out, phi0, S0, ra, ctl, \
                               tiles, bd.gr, bd.list, bd.count);                                               \
        else if (B == 1)                                                                                       \
            hipLaunchKernelGGL((k_redis3d_band<T, S>), dim3(BAND_GRID), dim3(NT), 0, st, phi, phi_out, phi0,   \
                               S0, ra, ctl, bd.gr, bd.list, bd.count, bd.next);                                \
    } while (0)
#define LSMP_R3K(S, B)                                                                                         \
    do {                                                                                                       \
        if (tiles) LSMP_R3(S, B, RKC_NB);                                                                      \
        else LSMP_R3(S, B, RKC);                                                                               \
    } while (0)
    if (btn) { if (stage == 1) LSMP_R3K(1, 1); else if (stage == 2) LSMP_R3K(2, 1); else LSMP_R3K(3, 1); }
    else { if (stage == 1) LSMP_R3K(1, 0); else if (stage == 2) LSMP_R3K(2, 0); else LSMP_R3K(3, 0); }
#undef LSMP_R3K
#undef LSMP_R3
}

template void launch_redis3d<double>(hipStream_t, int, int, const double *, double *, const double *, const double *,
                                     const RedisArgs &, unsigned long long *, const unsigned char *,
                                     const RedisBand<double> &, bool);
template void launch_redis3d<float>(hipStream_t, int, int, const float *, float *, const float *, const float *,
                                    const RedisArgs &, unsigned long long *, const unsigned char *,
                                    const RedisBand<float> &, bool);

}  // namespace lsmp
