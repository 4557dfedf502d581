// lsmp_adv_tile.hip -- the CRWENO5 line solve of an advection pass
// (level_set_rk3_solver.f90:1-221: crweno_flux_split :347-414, the WENO-JS
// boundary faces :288-345, the RK stage :16-88) for many lines of 34..513
// cells, or lines of 514..1025 cells, through LDS tiles.
//
// A workgroup of TW waves takes TL lines with neighbouring indices (for y and
// z lines, neighbouring in memory): 4 waves and 8 lines for lines of up to 513
// cells, 8 waves and 4 lines (one system per wave, 16 rows per lane) for lines
// of 514..1025 cells, which come in fewer lines per pass.
// (1) Every thread loads cells of the TL lines, coalesced -- along the line
// for x lines, across the lines for y and z lines -- and stores their split
// fluxes g and f in two LDS tiles, one row of slots per line, f in reversed
// cell order; then 4 TL threads form the WENO-JS boundary faces of the TL
// lines' two systems.  Reversing a line
// mirrors the stencil and exchanges A and C, which is what an f row does to a
// g row, so the f system of a line is the g system of its reversed f fluxes,
// with the reversed orientation's sign on the flux differences.  (2) Each
// wave solves 2 TL / TW of the TL lines' 2 TL systems in turn (g, and f
// reversed, both with the g rows): lane s eliminates its K = ceil((n-1)/64)
// rows with the two neighbouring interface values as parameters,
// X_i = p_i + q_i L + r_i R; the interface rows form a tridiagonal system
// across the 64 lanes, solved by parallel cyclic reduction; then each lane
// back-substitutes its rows, reading its stencils from the tiles.  (64 lanes
// per system measured faster than 32 with both systems in one wave: fewer
// rows per lane, 106 VGPRs instead of 144.)  The flux differences of the g and f systems go
// back into the tiles, over the line's own fluxes, which only its wave reads.
// (3) Every thread forms the RHS or the RK stage of cells of the TL lines and
// stores them, coalesced as in (1).  Nothing goes to scratch memory: the HBM
// traffic is the pass's own fields.  In MODE 2 the in-place update of phi is
// safe: the stencils run along the lines, and a workgroup reads and writes
// only its own lines.
//
// Same rows as the reference's (an f row is the g row of the mirrored stencil
// with A and C exchanged); the solve differs from one Thomas sweep only in
// rounding.
#include <hip/hip_runtime.h>

#include "lsmp_device.h"
#include "lsmp_internal.h"

namespace lsmp {

namespace {

// LDS slot of cell q (0-based) of a line: a pad slot after every 8 cells, so
// that the lanes of a system, K cells apart, spread over the banks
LSMP_DEV int tpos(int q) { return q + (q >> 3); }

}  // namespace

// slots per tile row for lines of n cells (odd: the y and z loads write TL
// rows at once)
static int adv_tile_row_slots(int n)
{
    const int rs = n + ((n - 1) >> 3) + 1;
    return rs | 1;
}

// TW waves per workgroup, TL lines per workgroup: 2 TL / TW systems per wave
template <typename T, int MODE, int STAGE, int KR, bool TUBE, int TL, int TW>
__global__ __launch_bounds__(TW * 64) void k_adv_tile(T *phi, const T *__restrict__ vel, T *R, T *L1, T *L2,
                                                       int nlines, int n, int lw, long ostride, long istride,
                                                       long estride, T dd, T dt, const int *__restrict__ lines,
                                                       TubeMask tm, int rs)
{
    static_assert((2 * TL) % TW == 0 && (TL & (TL - 1)) == 0, "whole systems per wave, TL a power of two");
    extern __shared__ __align__(16) unsigned char adv_lds[];
    T *tg = reinterpret_cast<T *>(adv_lds);   // TL rows of rs slots: g fluxes, then g flux differences
    T *tf = tg + TL * rs;                     // the same for f, cells in reversed order
    __shared__ size_t sbase[TL];
    // boundary faces: [line][0] g face 0, [1] g face n, [2] f face n, [3] f face 0
    // (the f system's start and end in its reversed order)
    __shared__ T sbc[TL][4];
    const int L0 = xcd_block() * TL;
    const int tid = threadIdx.x;
    if (tid < TL) {
        // lines past the end shadow the last one and store nothing
        const int Lc = min(L0 + tid, nlines - 1);
        const int bl = lines ? lines[Lc] : Lc;
        sbase[tid] = (size_t)(bl / lw) * (size_t)ostride + (size_t)(bl % lw) * (size_t)istride;
    }
    __syncthreads();
    const int ne = TL * n;
    const bool along = estride == 1;   // x lines: cells of a line are contiguous
    auto cell_of = [&](int e, int &l, int &q) {
        if (along) {
            l = e / n;
            q = e - l * n;
        } else {
            q = e / TL;
            l = e - q * TL;
        }
    };
    // (1) split fluxes of the TL lines into the tiles
    for (int e = tid; e < ne; e += TW * 64) {
        int l, q;
        cell_of(e, l, q);
        const size_t a = sbase[l] + (size_t)q * (size_t)estride;
        T g, f;
        flux_split(phi[a], vel[a], g, f);
        tg[l * rs + tpos(q)] = g;
        tf[l * rs + tpos(n - 1 - q)] = f;
    }
    __syncthreads();
    if (tid < 4 * TL) {
        // WENO-JS closures (crweno_flux_split :383-384, 409-410) on the clamped
        // cells in natural order: g on cells -2..2 or n-2..n+2, f on -1..3 or
        // n-1..n+3 (1-based)
        const int l = tid >> 2, w = tid & 3;
        const bool fs = w >= 2, right = (w == 1) || (w == 2);
        T v[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            const int q = clampi((right ? n - 2 : -2) + (fs ? 1 : 0) + t - 1, 0, n - 1);
            v[t] = fs ? tf[l * rs + tpos(n - 1 - q)] : tg[l * rs + tpos(q)];
        }
        sbc[l][w] = fs ? wenojs_f(v[0], v[1], v[2], v[3], v[4]) : wenojs_g(v[0], v[1], v[2], v[3], v[4]);
    }
    __syncthreads();

    // (2) one (line, system) at a time per wave, 64 lanes per system
    const int wave = tid >> 6, lane = tid & 63;
    const int s = lane;
    const int N = n - 1, K = (N + 63) / 64;
    const int lo = s * K + 1;
    const bool active = lo <= N;
    const bool last = (s + 1) * K >= N;
    const int hi = last ? N : (s + 1) * K - 1;
    const int nrow = active ? hi - lo + 1 : 0;
    const bool irow = active && !last;
    const bool next_last = (s + 2) * K >= N;
    for (int job = wave; job < 2 * TL; job += TW) {
        const int l = job >> 1;
        const bool isf = job & 1;
        const T sdd = isf ? -(T(1) / dd) : T(1) / dd;
        T *tile = (isf ? tf : tg) + l * rs;
        auto flux = [&](int c) { return tile[tpos(clampi(c - 1, 0, n - 1))]; };   // cell c = 1..n, clamped
        // fl[t] = flux at cell lo - 2 + t: the stencils of rows lo .. lo + KR
        T fl[KR + 6];
#pragma unroll
        for (int t = 0; t < KR + 6; ++t) fl[t] = flux(lo - 2 + t);
        auto coeffs = [&](int j, T &A, T &B, T &C, T &S) {   // row lo + j
            crweno_g_row_fast(fl[j], fl[j + 1], fl[j + 2], fl[j + 3], fl[j + 4], A, B, C, S);
        };
        // boundary faces at the system's start and end
        const T bc0 = sbc[l][isf ? 2 : 0], bcN = sbc[l][isf ? 3 : 1];
        // local sweep: X_i = d'_i + e'_i L - c'_i X_{i+1}; the row after the
        // last (j == nrow) is the interface row of a lane that has one
        T cp[KR], dp[KR], ep[KR];
        T A = T(0), B = T(1), C = T(0), S = T(0);   // interface row
#pragma unroll
        for (int j = 0; j < KR; ++j) {
            if (j < nrow || (irow && j == nrow)) {
                T a_, b_, c_, s_;
                coeffs(j, a_, b_, c_, s_);
                if (j == nrow) {
                    A = a_;
                    B = b_;
                    C = c_;
                    S = s_;
                } else if (j == 0) {
                    const T r = rcp_pos(b_);
                    cp[j] = c_ * r;
                    dp[j] = s_ * r;
                    ep[j] = -a_ * r;
                } else {
                    const T r = rcp_pos(b_ - a_ * cp[j - 1]);
                    cp[j] = c_ * r;
                    dp[j] = (s_ - a_ * dp[j - 1]) * r;
                    ep[j] = -a_ * ep[j - 1] * r;
                }
            }
        }
        // (p, q, r) of the last row and of the first
        T ph = T(0), qh = T(1), rh = T(0), pl = T(0), ql = T(0), rl = T(1);
#pragma unroll
        for (int j = KR - 1; j >= 0; --j) {
            if (j < nrow) {
                if (j == nrow - 1) {
                    pl = dp[j];
                    ql = ep[j];
                    rl = -cp[j];
                    ph = pl;
                    qh = ql;
                    rh = rl;
                } else {
                    pl = dp[j] - cp[j] * pl;
                    ql = ep[j] - cp[j] * ql;
                    rl = -cp[j] * rl;
                }
            }
        }
        // interface row (s+1)K: X_hi of this lane and X_lo of the next enter it
        T ra = T(0), rb = T(1), rc = T(0), rd = T(0);
        const T npl = __shfl(pl, (lane + 1) & 63, 64), nql = __shfl(ql, (lane + 1) & 63, 64);
        const T nrl = __shfl(rl, (lane + 1) & 63, 64), nbcN = bcN;
        if (irow) {
            ra = A * qh;
            rb = B + A * rh + C * nql;
            rc = C * nrl;
            rd = S - A * ph - C * npl;
            if (s == 0) {
                rd = rd - ra * bc0;
                ra = T(0);
            }
            if (next_last) {
                rd = rd - rc * nbcN;
                rc = T(0);
            }
        }
        // parallel cyclic reduction over the 64 lanes; each lane
        // hands its neighbours 1 / b with its row.  Past the ends of the line
        // the neighbour is an identity row: its multiplier is zero (the lane
        // then reads its own, finite row)
#pragma unroll
        for (int st = 1; st < 64; st <<= 1) {
            const bool hasm = s - st >= 0, hasp = s + st < 64;
            const int lm = hasm ? lane - st : lane, lp = hasp ? lane + st : lane;
            const T ib = rcp_w(rb);
            const T am = __shfl(ra, lm, 64), im = __shfl(ib, lm, 64), cm = __shfl(rc, lm, 64), dm = __shfl(rd, lm, 64);
            const T ap = __shfl(ra, lp, 64), ip = __shfl(ib, lp, 64), cq = __shfl(rc, lp, 64), dq = __shfl(rd, lp, 64);
            const T k1 = hasm ? ra * im : T(0), k2 = hasp ? rc * ip : T(0);
            const T na = -am * k1, nc = -cq * k2;
            const T nb = rb - cm * k1 - ap * k2, nd = rd - dm * k1 - dq * k2;
            ra = na;
            rb = nb;
            rc = nc;
            rd = nd;
        }
        const T Xb = rd * rcp_w(rb);
        const T Xprev = __shfl(Xb, (lane - 1) & 63, 64);
        const T Lv = s == 0 ? bc0 : Xprev;
        const T Rv = last ? bcN : Xb;

        // back-substitution: the flux differences of cells lo .. hi + 1 of this
        // system go over the line's fluxes (every lane of the wave has read its
        // stencils above)
        T x = Rv;
#pragma unroll
        for (int j = KR - 1; j >= 0; --j) {
            if (j < nrow) {
                const T xn = dp[j] + ep[j] * Lv - cp[j] * x;
                tile[tpos(lo + j)] = (xn - x) * sdd;   // cell lo + j + 1
                x = xn;
            }
        }
        if (active) tile[tpos(lo - 1)] = (Lv - x) * sdd;   // cell lo
    }
    __syncthreads();

    // (3) the RHS or the RK stage of every cell of the TL lines
    for (int e = tid; e < ne; e += TW * 64) {
        int l, q;
        cell_of(e, l, q);
        if (L0 + l >= nlines) continue;
        const size_t a = sbase[l] + (size_t)q * (size_t)estride;
        const T t1 = tg[l * rs + tpos(q)], t2 = tf[l * rs + tpos(n - 1 - q)];
        if (MODE == 0) {
            R[a] = t1 + t2;
        } else if (MODE == 1) {
            R[a] = (R[a] + t1) + t2;
        } else if (!TUBE || tm.in(a)) {
            const T r = (R[a] + t1) + t2, Pv = phi[a];
            if (STAGE == 1) {
                L1[a] = r;
                phi[a] = Pv + r * dt;
            } else if (STAGE == 2) {
                L2[a] = r;
                phi[a] = Pv + (T(-3) * L1[a] + r) / T(4) * dt;
            } else {
                phi[a] = Pv + (-L1[a] - L2[a] + T(8) * r) / T(12) * dt;
            }
        }
    }
}

template <typename T>
void launch_adv_tile(hipStream_t st, int mode, int stage, T *phi, const T *vel, T *R, T *L1, T *L2, int nlines, int n,
                     int lw, long ostride, long istride, long estride, T dd, T dt, const int *lines,
                     const TubeMask &tm)
{
    const int rs = adv_tile_row_slots(n);
    const bool tube = tm.fx != nullptr;
    const bool k4 = n <= 64 * 4 + 1;   // four rows per lane suffice, else eight (up to 513 cells)
    // lines of 514..1025 cells: sixteen rows per lane, four lines per workgroup
    const bool k16 = n > 64 * 8 + 1;
    const int tl = k16 ? 4 : 8;
    const size_t lds = 2 * (size_t)tl * rs * sizeof(T);
    const dim3 g((unsigned)((nlines + tl - 1) / tl)), b(k16 ? 8 * 64 : 4 * 64);
    // tiles of lines longer than 257 cells take more than 64 KiB of LDS (FP64)
#define LSMP_TILEK(M, S, KR, TB, TLV, TWV)                                                                     \
    do {                                                                                                       \
        if (lds > 65536)                                                                                       \
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_adv_tile<T, M, S, KR, TB, TLV, TWV>),  \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                   \
        hipLaunchKernelGGL((k_adv_tile<T, M, S, KR, TB, TLV, TWV>), g, b, lds, st, phi, vel, R, L1, L2, nlines, \
                           n, lw, ostride, istride, estride, dd, dt, lines, tm, rs);                           \
    } while (0)
#define LSMP_TILE(M, S)                                                                                        \
    do {                                                                                                       \
        if (k16) LSMP_TILEK(M, S, 16, false, 4, 8);                                                            \
        else if constexpr (M == 2) {                                                                           \
            if (k4 && tube) LSMP_TILEK(M, S, 4, true, 8, 4);                                                   \
            else if (k4) LSMP_TILEK(M, S, 4, false, 8, 4);                                                     \
            else if (tube) LSMP_TILEK(M, S, 8, true, 8, 4);                                                    \
            else LSMP_TILEK(M, S, 8, false, 8, 4);                                                             \
        } else {                                                                                               \
            if (k4) LSMP_TILEK(M, S, 4, false, 8, 4);                                                          \
            else LSMP_TILEK(M, S, 8, false, 8, 4);                                                             \
        }                                                                                                      \
    } while (0)
    if (mode == 0) LSMP_TILE(0, 0);
    else if (mode == 1) LSMP_TILE(1, 1);
    else if (stage == 1) LSMP_TILE(2, 1);
    else if (stage == 2) LSMP_TILE(2, 2);
    else LSMP_TILE(2, 3);
#undef LSMP_TILE
#undef LSMP_TILEK
}

#define LSMP_INST_TILE(T)                                                                                      \
    template void launch_adv_tile<T>(hipStream_t, int, int, T *, const T *, T *, T *, T *, int, int, int, long, \
                                     long, long, T, T, const int *, const TubeMask &);
LSMP_INST_TILE(double)
LSMP_INST_TILE(float)
#undef LSMP_INST_TILE

}  // namespace lsmp
