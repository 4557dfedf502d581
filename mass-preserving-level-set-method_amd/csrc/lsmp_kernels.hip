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

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

// ---------------------------------------------------------------------------
// Advection (level_set_rk3_solver.f90:1-221): CRWENO5 line solves
// (crweno_flux_split :347-414, wenojs_flux_split :288-345, Thomas
// TDMA.f90:1-36) along x, y and z, accumulated into the RHS; the last
// direction applies the TVD-RK3 stage (:16-88) in place.
//
// Fortran cells c = 1..n of a line map to 0-based q = c-1; interface m
// (0..n) is the face between cells m and m+1.  The two boundary faces come
// from WENO-JS, faces 1..n-1 from the two tridiagonal systems.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Shared by both line-solve layouts: an f row is the g row of the mirrored
// stencil f(i+3..i-1) with A and C exchanged (crweno_flux_split :388-407
// against :362-381), so lanes solving g and f systems run one instruction
// stream; the two boundary faces come from WENO-JS per system (:383-384,
// 409-410).  Each lane runs the reference's Thomas recurrence (TDMA.f90:1-36)
// on its rows: the forward sweep keeps (c', d') in
// scr[((row-1)*2 + comp)*nslot + slot], slot = 64 w + lane (coalesced), and
// the backward sweep forms the flux difference of each cell; the f lane hands
// its half to the g lane by a cross-half shuffle (lane ^ 32) and the g lane
// applies the RHS or the RK stage.  Loads are prefetched into registers a
// chunk of rows ahead (AXB rows of scratch in the x kernel's backward sweep).
//
// Line L starts at (L / lw) * ostride + (L % lw) * istride with element
// stride estride (x lines: lw = nlines, istride = nx, estride = 1).
// ---------------------------------------------------------------------------
constexpr int AXB = 8;

// ---------------------------------------------------------------------------
// Twisted line solve: four lanes per line.  The rows 1..n-1 of each system
// are eliminated from both ends at once and meet at the twist face k = n/2:
// lanes 0-15 of a wave take the g system of lines 16 w + 0..15 from face 1
// upward, lanes 16-31 the same g systems from face n-1 downward, lanes 32-47
// and 48-63 the f systems likewise.  A lane that starts at the far end runs
// the same code on the reversed line (primed cell c' = n + 1 - c): reversing
// a line mirrors the stencil and exchanges A and C, which is what the f rows
// already do, so every lane computes its rows as crweno_g_row of the window
// in its own order, mirrored when isf != bot.  Each lane runs the Thomas
// forward recurrence (TDMA.f90:1-36) over its len rows into scratch; the two
// halves then solve the twist row together,
//   (B - A c'_top - C c'_bot) X_k = S - A d'_top - C d'_bot,
// and each back-substitutes over its own rows from X_k, emitting the flux
// difference of every cell (the far-end lanes with the sign of the reversed
// orientation).  Same arithmetic per row as one Thomas sweep, half the
// sequential length, twice the lanes.  With twice the waves resident, the
// rolling prefetch is shallower than the two-lane kernel's (TXC rows ahead in
// the forward sweep, TXB in the backward one), which keeps the VGPRs low
// enough for more waves per SIMD.
// ---------------------------------------------------------------------------
// rows of forward / backward prefetch: FP64 4 / 4 (about 100 VGPRs, four
// waves per SIMD), FP32 8 / 4 (five waves); measured best at 256^3
template <typename T> struct TwistDepth { static constexpr int C = 8, B = 4; };
template <> struct TwistDepth<double> { static constexpr int C = 4, B = 4; };

template <typename T, int MODE, int STAGE, bool TUBE>
__global__ __launch_bounds__(256) void k_adv_twist(T *phi, const T *__restrict__ vel, T *R, T *L1, T *L2,
                                                    T *__restrict__ scr, int nlines, int n, int lw, long ostride,
                                                    long istride, long estride, T dd, T dt,
                                                    const int *__restrict__ lines, TubeMask tm)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int nwaves = (nlines + 15) >> 4;
    if (wave >= nwaves) return;   // whole waves only: the shuffles below need every lane
    constexpr int TXC = TwistDepth<T>::C, TXB = TwistDepth<T>::B;
    const bool isf = lane >= 32, bot = (lane & 16) != 0, mir = isf != bot;
    const int L = wave * 16 + (lane & 15);
    const bool act = L < nlines;
    const int Lc = act ? L : nlines - 1;   // idle lanes shadow the last line and store nothing
    const int bl = lines ? lines[Lc] : Lc;  // with a list, lane group L solves listed line L
    const size_t base = (size_t)(bl / lw) * (size_t)ostride + (size_t)(bl % lw) * (size_t)istride;
    const size_t nslot = (size_t)nwaves * 64, slot = (size_t)wave * 64 + lane;
    const int k = n / 2;                         // twist face
    const int len = bot ? n - 1 - k : k - 1;     // rows this lane eliminates (>= 1 for n >= 4)
    const T inv_dd = T(1) / dd;
    const T sdd = bot ? -inv_dd : inv_dd;        // flux difference in this lane's orientation
    // this lane's cell c' = 1..n -> element of the line, clamped (zero-gradient ghosts)
    auto el = [&](int c) -> int { return clampi(bot ? n - c : c - 1, 0, n - 1); };
    auto at = [&](int c) -> size_t { return base + (size_t)el(c) * (size_t)estride; };
    auto flux = [&](T p, T u) {
        T g, f;
        flux_split(p, u, g, f);
        return isf ? f : g;
    };
    auto ld = [&](int c) {
        const size_t a = at(c);
        return flux(phi[a], vel[a]);
    };
    auto emit = [&](size_t a, T t1, T t2, T Rv, T Pv, T L1v, T L2v) {
        if (MODE == 1) {
            R[a] = (Rv + t1) + t2;
        } else if (!TUBE || tm.in(a)) {
            const T r = (Rv + t1) + t2;
            if (STAGE == 1) {
                L1[a] = r;
                phi[a] = Pv + r * dt;
            } else if (STAGE == 2) {
                L2[a] = r;
                phi[a] = Pv + (T(-3) * L1v + r) / T(4) * dt;
            } else {
                phi[a] = Pv + (-L1v - L2v + T(8) * r) / T(12) * dt;
            }
        }
    };

    // boundary face where this lane starts: face 0 (top) or face n (bottom),
    // from the WENO-JS closure of its system on the cells in natural order
    T bc;
    {
        T v[6], o[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) v[t] = ld(-2 + t);
#pragma unroll
        for (int t = 0; t < 6; ++t) o[t] = bot ? v[5 - t] : v[t];
        bc = isf ? wenojs_f(o[1], o[2], o[3], o[4], o[5]) : wenojs_g(o[0], o[1], o[2], o[3], o[4]);
    }

    // forward sweep over rows j = 1..len; w[t] holds the flux at cell j - 2 + t
    T w[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) w[t] = ld(-1 + t);
    auto coeffs = [&](T &A, T &B, T &C, T &S) {
        const T x0 = mir ? w[5] : w[0], x1 = mir ? w[4] : w[1], x2 = mir ? w[3] : w[2];
        const T x3 = mir ? w[2] : w[3], x4 = mir ? w[1] : w[4];
        T Ag, Cg;
        crweno_g_row_fast(x0, x1, x2, x3, x4, Ag, B, Cg, S);
        A = mir ? Cg : Ag;
        C = mir ? Ag : Cg;
    };
    auto shift = [&](T next) {
#pragma unroll
        for (int t = 0; t < 5; ++t) w[t] = w[t + 1];
        w[5] = next;
    };
    T cp, dp;
    size_t o = slot;   // scratch offset of the current row: (j - 1) * 2 * nslot + slot
    {
        T A, B, C, S;
        coeffs(A, B, C, S);
        S = S - A * bc;
        const T rr = rcp_pos(B);
        cp = C * rr;
        dp = S * rr;
        if (act) {
            scr[o] = cp;
            scr[o + nslot] = dp;
        }
        o += 2 * nslot;
        shift(ld(5));
    }
    auto mid = [&](T next) {
        T A, B, C, S;
        coeffs(A, B, C, S);
        const T rr = rcp_pos(B - A * cp);
        const T c = C * rr;
        const T d = (S - A * dp) * rr;
        if (act) {
            scr[o] = c;
            scr[o + nslot] = d;
        }
        o += 2 * nslot;
        cp = c;
        dp = d;
        shift(next);
    };
    // row j consumes cell j + 4; slot r of the rolling buffer serves rows
    // j = 2 + r (mod TXC) and is reloaded TXC rows ahead as soon as it is used
    T pp[TXC], pu[TXC];
    int m0 = 2;
    if (m0 + TXC - 1 <= len) {
#pragma unroll
        for (int r = 0; r < TXC; ++r) {
            const size_t ad = at(m0 + 4 + r);
            pp[r] = phi[ad];
            pu[r] = vel[ad];
        }
    }
    for (; m0 + TXC - 1 <= len; m0 += TXC) {
#pragma unroll
        for (int r = 0; r < TXC; ++r) {
            const T next = flux(pp[r], pu[r]);
            const size_t ad = at(m0 + TXC + 4 + r);   // clamped: reads past the line end are harmless
            pp[r] = phi[ad];
            pu[r] = vel[ad];
            mid(next);
        }
    }
    for (int j = m0; j <= len; ++j) mid(ld(j + 4));

    // twist row: face len + 1 of this lane = face k of the line
    T X;
    {
        T A, B, C, S;
        coeffs(A, B, C, S);
        const T oc = __shfl_xor(cp, 16, 64), od = __shfl_xor(dp, 16, 64);
        const T Xk = (S - A * dp - C * od) * rcp_pos(B - A * cp - C * oc);
        X = __shfl(Xk, lane & ~16, 64);   // both halves continue from the top lane's value
    }

    // back-substitution: X(j) = d'(j) - c'(j) X(j+1) for j = len .. 1,
    // emitting cell j + 1 of this lane (between its faces j and j + 1)
    long ae = (long)(base + (size_t)el(len + 1) * (size_t)estride);
    const long astep = bot ? estride : -estride;
    auto back = [&](T c, T d, T Rv, T Pv, T L1v, T L2v) {
        const T X1 = d - c * X;
        const T t = (X1 - X) * sdd;
        const T t2 = __shfl_xor(t, 32, 64);
        if (!isf && act) emit((size_t)ae, t, t2, Rv, Pv, L1v, L2v);
        ae += astep;
        X = X1;
    };
    // rolling buffer of the scratch rows and RHS operands, TXB rows ahead
    T pc[TXB], pd[TXB], pR[TXB], pP[TXB], pL1[TXB], pL2[TXB];
    auto load_b = [&](int r, int j) {   // row j >= 1 into slot r
        const size_t ob = (size_t)(j - 1) * 2 * nslot + slot;
        pc[r] = scr[ob];
        pd[r] = scr[ob + nslot];
        if (!isf) {
            const size_t ad = at(j + 1);
            pR[r] = R[ad];
            if (MODE == 2) pP[r] = phi[ad];
            if (MODE == 2 && STAGE >= 2) pL1[r] = L1[ad];
            if (MODE == 2 && STAGE == 3) pL2[r] = L2[ad];
        }
    };
    m0 = len;
    if (m0 - TXB + 1 >= 1) {
#pragma unroll
        for (int r = 0; r < TXB; ++r) load_b(r, m0 - r);
    }
    for (; m0 - TXB + 1 >= 1; m0 -= TXB) {
        const bool more = m0 - 2 * TXB + 1 >= 1;   // the next chunk is full: reload
#pragma unroll
        for (int r = 0; r < TXB; ++r) {
            const T c = pc[r], d = pd[r], Rv = pR[r], Pv = pP[r], L1v = pL1[r], L2v = pL2[r];
            if (more) load_b(r, m0 - TXB - r);
            back(c, d, Rv, Pv, L1v, L2v);
        }
    }
    for (int j = m0; j >= 1; --j) {   // tail
        const size_t ob = (size_t)(j - 1) * 2 * nslot + slot;
        T Rv = T(0), Pv = T(0), L1v = T(0), L2v = T(0);
        if (!isf) {
            Rv = R[(size_t)ae];
            if (MODE == 2) Pv = phi[(size_t)ae];
            if (MODE == 2 && STAGE >= 2) L1v = L1[(size_t)ae];
            if (MODE == 2 && STAGE == 3) L2v = L2[(size_t)ae];
        }
        back(scr[ob], scr[ob + nslot], Rv, Pv, L1v, L2v);
    }
    // the boundary cell: between the boundary face and face 1 of this lane
    {
        const T t = (bc - X) * sdd;
        const T t2 = __shfl_xor(t, 32, 64);
        if (!isf && act) {
            const size_t a = at(1);
            T L1v = T(0), L2v = T(0);
            const T Rv = R[a];
            const T Pv = MODE == 2 ? phi[a] : T(0);
            if (MODE == 2 && STAGE >= 2) L1v = L1[a];
            if (MODE == 2 && STAGE == 3) L2v = L2[a];
            emit(a, t, t2, Rv, Pv, L1v, L2v);
        }
    }
}

// Paired solve for lines along x (contiguous in memory), first direction
// only (R = t1 + t2): two lanes per line, lane l < 32 of a wave on the g
// system of line 32 w + l, lane l + 32 on its f system, each over all rows
// 1..n-1.  (A twisted x kernel measured no faster: these lines are already
// read coalesced through LDS.)  Per wave, the 32 lines' cells arrive in
// chunks of XW through an LDS tile that the wave loads cooperatively (16
// consecutive elements per 128-byte segment, coalesced); the next chunk is
// loaded into registers while the current one is consumed.  The backward
// sweep writes its flux differences into the same tile and flushes each
// chunk with coalesced stores.  Every wave of the workgroup runs the same
// chunk loop (idle waves shadow a valid line and store nothing), so the
// workgroup barriers are uniform.
constexpr int XW = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_adv_pair_x(const T *__restrict__ phi, const T *__restrict__ vel,
                                                     T *__restrict__ R, T *__restrict__ scr, int nlines, int n,
                                                     T dd)
{
    __shared__ T tp[4][32][XW + 1];
    __shared__ T tu[4][32][XW + 1];
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, wv = threadIdx.x >> 6;
    const int nwaves = (nlines + 31) >> 5;
    const bool isf = lane >= 32;
    const int ln = lane & 31;
    const int L = wave * 32 + ln;
    const bool act = L < nlines;
    const size_t base = (size_t)(act ? L : nlines - 1) * (size_t)n;
    const size_t nslot = (size_t)nwaves * 64;
    const size_t slot = (size_t)(wave < nwaves ? wave : nwaves - 1) * 64 + lane;
    const T inv_dd = T(1) / dd;
    auto at = [&](int c) -> size_t { return base + (size_t)clampi(c - 1, 0, n - 1); };
    auto flux = [&](T p, T u) {
        T g, f;
        flux_split(p, u, g, f);
        return isf ? f : g;
    };
    auto ld = [&](int c) {
        const size_t a = at(c);
        return flux(phi[a], vel[a]);
    };
    // cooperative chunk access: element e of line l of chunk k is item
    // lane + 64 i, l = item / XW, e = item % XW
    constexpr int PER = 32 * XW / 64;
    auto chunk_addr = [&](int k, int i, int &l, int &e) -> size_t {
        const int item = lane + 64 * i;
        l = item / XW;
        e = item % XW;
        int line = wave * 32 + l;
        line = line < nlines ? line : nlines - 1;
        return (size_t)line * (size_t)n + (size_t)clampi(k * XW + e, 0, n - 1);
    };
    T rp[PER], ru[PER];
    auto load_regs = [&](int k) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int l, e;
            const size_t a = chunk_addr(k, i, l, e);
            rp[i] = phi[a];
            ru[i] = vel[a];
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int item = lane + 64 * i;
            tp[wv][item / XW][item % XW] = rp[i];
            tu[wv][item / XW][item % XW] = ru[i];
        }
    };

    T bc0, bcN;
    {
        T v[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) v[t] = ld(-2 + t);
        bc0 = isf ? wenojs_f(v[1], v[2], v[3], v[4], v[5]) : wenojs_g(v[0], v[1], v[2], v[3], v[4]);
#pragma unroll
        for (int t = 0; t < 6; ++t) v[t] = ld(n - 2 + t);
        bcN = isf ? wenojs_f(v[1], v[2], v[3], v[4], v[5]) : wenojs_g(v[0], v[1], v[2], v[3], v[4]);
    }
    T w[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) w[t] = ld(-1 + t);
    auto coeffs = [&](T &A, T &B, T &C, T &S) {
        const T x0 = isf ? w[5] : w[0], x1 = isf ? w[4] : w[1], x2 = isf ? w[3] : w[2];
        const T x3 = isf ? w[2] : w[3], x4 = isf ? w[1] : w[4];
        T Ag, Cg;
        crweno_g_row_fast(x0, x1, x2, x3, x4, Ag, B, Cg, S);
        A = isf ? Cg : Ag;
        C = isf ? Ag : Cg;
    };
    auto shift = [&](T next) {
#pragma unroll
        for (int t = 0; t < 5; ++t) w[t] = w[t + 1];
        w[5] = next;
    };
    T cp, dp;
    size_t o = slot;
    {
        T A, B, C, S;
        coeffs(A, B, C, S);
        S = S - A * bc0;
        const T rr = rcp_pos(B);
        cp = C * rr;
        dp = S * rr;
        if (act) {
            scr[o] = cp;
            scr[o + nslot] = dp;
        }
        o += 2 * nslot;
        shift(ld(5));
    }
    auto mid = [&](T next) {
        T A, B, C, S;
        coeffs(A, B, C, S);
        const T rr = rcp_pos(B - A * cp);
        const T c = C * rr;
        const T d = (S - A * dp) * rr;
        if (act) {
            scr[o] = c;
            scr[o + nslot] = d;
        }
        o += 2 * nslot;
        cp = c;
        dp = d;
        shift(next);
    };
    // rows 2 .. n-2; row m consumes cell m + 4 = element m + 3 of chunk (m + 3) / XW
    const int kfirst = 5 / XW, klast = (n + 1) / XW;
    load_regs(kfirst);
    for (int k = kfirst; k <= klast; ++k) {
        __syncthreads();   // every wave is done with the previous chunk
        store_tile();
        __syncthreads();
        if (k < klast) load_regs(k + 1);
        const int mlo = max(2, k * XW - 3), mhi = min(n - 2, k * XW + XW - 4);
        for (int m = mlo; m <= mhi; ++m) {
            const int e = m + 3 - k * XW;
            mid(flux(tp[wv][ln][e], tu[wv][ln][e]));
        }
    }
    {
        T A, B, C, S;
        coeffs(A, B, C, S);
        S = S - C * bcN;
        dp = (S - A * dp) * rcp_pos(B - A * cp);
    }

    // backward sweep; cell c = m + 1 goes to element m of the output tile
    T X = dp;
    auto put = [&](int q, T t) {   // flux difference t of this lane's system at element q
        const T t2 = __shfl_xor(t, 32, 64);
        if (!isf) tp[wv][ln][q % XW] = t + t2;
    };
    auto flush = [&](int k) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int l, e;
            chunk_addr(k, i, l, e);
            const int q = k * XW + e;
            if (wave * 32 + l < nlines && q < n) R[(size_t)(wave * 32 + l) * (size_t)n + q] = tp[wv][l][e];
        }
        __syncthreads();
    };
    __syncthreads();   // the forward sweep's last tile reads are done
    int kout = (n - 1) / XW;
    put(n - 1, (X - bcN) * inv_dd);
    T pc[AXB], pd[AXB];
    auto load_b = [&](int r, int m) {
        const size_t ob = (size_t)(m - 1) * 2 * nslot + slot;
        pc[r] = scr[ob];
        pd[r] = scr[ob + nslot];
    };
    auto back = [&](int m, T c, T d) {
        if (m / XW != kout) {   // element m starts a lower chunk: write the finished one
            flush(kout);
            kout = m / XW;
        }
        const T X1 = d - c * X;
        put(m, (X1 - X) * inv_dd);
        X = X1;
    };
    int m0 = n - 2;
    if (m0 - AXB + 1 >= 1) {
#pragma unroll
        for (int r = 0; r < AXB; ++r) load_b(r, m0 - r);
    }
    for (; m0 - AXB + 1 >= 1; m0 -= AXB) {
        const bool more = m0 - 2 * AXB + 1 >= 1;
#pragma unroll
        for (int r = 0; r < AXB; ++r) {
            const T c = pc[r], d = pd[r];
            if (more) load_b(r, m0 - AXB - r);
            back(m0 - r, c, d);
        }
    }
    for (int m = m0; m >= 1; --m) {
        const size_t ob = (size_t)(m - 1) * 2 * nslot + slot;
        back(m, scr[ob], scr[ob + nslot]);
    }
    if (kout != 0) {
        flush(kout);
        kout = 0;
    }
    put(0, (bc0 - X) * inv_dd);
    flush(0);
}

// Partitioned solve for lines of 34..257 cells: one wave per line, lanes
// 0-31 on its g system and 32-63 on its f system; MODE and STAGE as in
// k_adv_twist (0: R = t1 + t2, 1: R += t1 + t2, 2: RK stage in place).  With K = ceil((n-1)/32) (2..8), lane s of a half owns the
// rows lo = sK+1 .. hi = (s+1)K-1 and the interface row (s+1)K; the last
// lane owns the rows up to n-1 and no interface row.  Each lane runs the
// Thomas recurrence (TDMA.f90:1-36) over its rows with the two neighbouring
// interface values L and R as parameters, X_i = p_i + q_i L + r_i R, keeping
// (c', d', e') of its rows in registers; every interface row then becomes one
// row of a tridiagonal system in the interface values, one per lane, solved
// by parallel cyclic reduction across the 32 lanes of the half (identity rows
// stand in beyond the last interface); each lane back-substitutes its rows
// from L and R and emits the flux differences of cells lo..hi+1.  Nothing
// goes to scratch memory.  Same rows as the reference's (an f row is the g
// row of the mirrored stencil with A and C exchanged); the solve differs from
// one Thomas sweep only in rounding.  In MODE 2 the in-place update is safe:
// a line's cells are read by its own wave only, all before the reduction.

template <typename T, int MODE, int STAGE, int KR, bool TUBE>
__global__ __launch_bounds__(256) void k_adv_part(T *phi, const T *__restrict__ vel, T *R, T *L1, T *L2,
                                                   int nlines, int n, int lw, long ostride, long istride,
                                                   long estride, T dd, T dt, const int *__restrict__ lines,
                                                   TubeMask tm)
{
    const int tid = xcd_block() * blockDim.x + threadIdx.x;
    const int wline = tid >> 6, lane = tid & 63;
    if (wline >= nlines) return;   // a whole wave per line
    const int line = lines ? lines[wline] : wline;   // with a list, wave w solves listed line w
    const bool isf = lane >= 32;
    const int s = lane & 31;
    const int N = n - 1, K = (N + 31) / 32;
    const int lo = s * K + 1;
    const bool active = lo <= N;
    const bool last = (s + 1) * K >= N;          // no interface row: the right value is face n
    const int hi = last ? N : (s + 1) * K - 1;
    const int nrow = active ? hi - lo + 1 : 0;   // 1..K
    const size_t base = (size_t)(line / lw) * (size_t)ostride + (size_t)(line % lw) * (size_t)istride;
    const T inv_dd = T(1) / dd;
    auto at = [&](int c) -> size_t { return base + (size_t)clampi(c - 1, 0, n - 1) * (size_t)estride; };
    auto emit = [&](int c, T t1, T t2) {   // cell c of the line
        const size_t a = at(c);
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
    };
    auto flux = [&](int c) {
        const size_t a = at(c);
        T g, f;
        flux_split(phi[a], vel[a], g, f);
        return isf ? f : g;
    };
    // fl[t] = flux at cell lo - 2 + t: the stencils of rows lo .. lo + KR (the
    // last one is the interface row of a full segment)
    T fl[KR + 6];
#pragma unroll
    for (int t = 0; t < KR + 6; ++t) fl[t] = flux(lo - 2 + t);
    auto coeffs = [&](int j, T &A, T &B, T &C, T &S) {   // row lo + j
        const T x0 = isf ? fl[j + 5] : fl[j], x1 = isf ? fl[j + 4] : fl[j + 1], x2 = isf ? fl[j + 3] : fl[j + 2];
        const T x3 = isf ? fl[j + 2] : fl[j + 3], x4 = isf ? fl[j + 1] : fl[j + 4];
        T Ag, Cg;
        crweno_g_row_fast(x0, x1, x2, x3, x4, Ag, B, Cg, S);
        A = isf ? Cg : Ag;
        C = isf ? Ag : Cg;
    };
    // boundary faces of this system, where a lane needs them
    T bc0 = T(0), bcN = T(0);
    if (s == 0) {
        T v[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) v[t] = flux(-2 + t);
        bc0 = isf ? wenojs_f(v[1], v[2], v[3], v[4], v[5]) : wenojs_g(v[0], v[1], v[2], v[3], v[4]);
    }
    if (active && last) {
        T v[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) v[t] = flux(n - 2 + t);
        bcN = isf ? wenojs_f(v[1], v[2], v[3], v[4], v[5]) : wenojs_g(v[0], v[1], v[2], v[3], v[4]);
    }

    // local sweep: X_i = d'_i + e'_i L - c'_i X_{i+1}; the row after the last
    // (j == nrow) is the interface row of a lane that has one
    const bool irow = active && !last;
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
    const T nrl = __shfl(rl, (lane + 1) & 63, 64), nbcN = __shfl(bcN, (lane + 1) & 63, 64);
    const bool next_last = (s + 2) * K >= N;
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
    // parallel cyclic reduction over the 32 lanes of the half
#pragma unroll
    for (int st = 1; st < 32; st <<= 1) {
        const bool hasm = s - st >= 0, hasp = s + st < 32;
        const int lm = hasm ? lane - st : lane, lp = hasp ? lane + st : lane;
        T am = __shfl(ra, lm, 64), bm = __shfl(rb, lm, 64), cm = __shfl(rc, lm, 64), dm = __shfl(rd, lm, 64);
        T ap = __shfl(ra, lp, 64), bp = __shfl(rb, lp, 64), cq = __shfl(rc, lp, 64), dq = __shfl(rd, lp, 64);
        if (!hasm) { am = T(0); bm = T(1); cm = T(0); dm = T(0); }
        if (!hasp) { ap = T(0); bp = T(1); cq = T(0); dq = T(0); }
        const T k1 = ra / bm, k2 = rc / bp;
        const T na = -am * k1, nc = -cq * k2;
        const T nb = rb - cm * k1 - ap * k2, nd = rd - dm * k1 - dq * k2;
        ra = na;
        rb = nb;
        rc = nc;
        rd = nd;
    }
    const T Xb = rd / rb;
    const T Xprev = __shfl(Xb, (lane - 1) & 63, 64);
    const T Lv = s == 0 ? bc0 : Xprev;
    const T Rv = last ? bcN : Xb;

    // back-substitution and the flux differences of cells lo .. hi + 1
    T x = Rv;
#pragma unroll
    for (int j = KR - 1; j >= 0; --j) {
        if (j < nrow) {
            const T xn = dp[j] + ep[j] * Lv - cp[j] * x;
            const T t = (xn - x) * inv_dd;
            const T t2 = __shfl_xor(t, 32, 64);
            if (!isf) emit(lo + j + 1, t, t2);
            x = xn;
        }
    }
    {
        const T t = (Lv - x) * inv_dd;
        const T t2 = __shfl_xor(t, 32, 64);
        if (!isf && active) emit(lo, t, t2);
    }
}


// ---------------------------------------------------------------------------
// Reinitialisation stage, 2D (level_set_rk3_redis.f90:176-264 with the
// gradient :266-369, lambda :371-454 and ls_funs ns_solver/others.f90:1-45);
// the 3D stage is in lsmp_redis.hip.
// ---------------------------------------------------------------------------

// Extended-tile position (row ey, column ex) of ring cell number u of a
// TX x TY tile + 1 ring: the top and bottom rows, then the left and right
// columns.
template <int TX, int TY> LSMP_DEV void ext_ring_pos(int u, int &ey, int &ex)
{
    constexpr int EX = TX + 2, EY = TY + 2;
    if (u < EX) { ey = 0; ex = u; }
    else if (u < 2 * EX) { ey = EY - 1; ex = u - EX; }
    else if (u < 2 * EX + TY) { ey = 1 + u - 2 * EX; ex = 0; }
    else { ey = 1 + u - 2 * EX - TY; ex = EX - 1; }
}

// Thread t's cell of the extended 2D tile: tile cells first, ring after them.
LSMP_DEV void ext_pos(int t, int &ey, int &ex)
{
    if (t < RTX * RTY) { ex = 1 + t % RTX; ey = 1 + t / RTX; }
    else if (t < REX * REY) ext_ring_pos<RTX, RTY>(t - RTX * RTY, ey, ex);
    else { ex = 0; ey = 0; }
}

// delta(phi) != 0 exactly where ls_funs_fast takes its middle branch
template <typename T> LSMP_DEV bool in_band(T x)
{
    return !(x > T(1) - T(1.0e-12)) && !(x < T(-1) + T(1.0e-12));
}

// 2D stage: one tile, no marching (Disk/ and Vortex2D/ forms: |grad| without
// z, lambda box 3x3 with centre weight 17, Disk/...redis.f90:357-361, 434-438).
// The tile's RTX x RTY cells and the ring around them (the lambda box) need
// D+ and D- in x and y.  Each WENO-HJ face m+1/2 gives the D+ term of cell m
// and the D- term of cell m+1 (weno_hj_face), so the workgroup forms every
// face once -- x faces i+1/2 for i = i0-2 .. i0+RTX on the REY ring rows, y
// faces j+1/2 for j = j0-2 .. j0+RTY on the REX ring columns, two faces a
// thread -- into LDS, and each cell of the tile and ring reads its four terms
// there.
constexpr int R2FX = (RTX + 3) * REY;   // x faces
constexpr int R2F = R2FX + REX * (RTY + 3);   // and y faces
static_assert(R2F <= 2 * RNT, "two faces a thread");

template <typename T, int STAGE, int BTN>
__global__ __launch_bounds__(RNT) void k_redis_stage2d(const T *__restrict__ phi, T *__restrict__ phi_out,
                                                       const T *__restrict__ phi0, const T *__restrict__ S0,
                                                       RedisArgs ra, unsigned long long *ctl)
{
    __shared__ T sphi[RPY][RPX];
    __shared__ T fplus[R2F], fminus[R2F];   // the faces' D+ and D- terms, times 12
    __shared__ T ss2[REY][REX];
    __shared__ T ss3[REY][REX];
    if (ctl[1]) return;   // the pseudo-time loop has stopped (k_redis_ctl)
    const int nx = ra.nx, ny = ra.ny;
    const T inv12dx = (T)ra.inv12dx, inv12dy = (T)ra.inv12dy;
    const T inv_w = (T)ra.inv_ls_wid, rdt = (T)ra.rdt, c3 = (T)ra.three;
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * RTX, j0 = blockIdx.y * RTY;
    const bool ext = t < REX * REY;
    int ex, ey;
    ext_pos(t, ey, ex);   // tile cells on whole waves, the ring after them
    const int gi = clampi(i0 - 1 + ex, 0, nx - 1), gj = clampi(j0 - 1 + ey, 0, ny - 1);
    const bool inner = t < RTX * RTY && (i0 - 1 + ex) < nx && (j0 - 1 + ey) < ny;
    const int pc = gi - (i0 - 4), pr = gj - (j0 - 4);
    for (int idx = t; idx < RPX * RPY; idx += RNT) {
        int c = idx % RPX, r = idx / RPX;
        int gx = clampi(i0 - 4 + c, 0, nx - 1), gy = clampi(j0 - 4 + r, 0, ny - 1);
        sphi[r][c] = phi[(size_t)gy * nx + gx];
    }
    const T s0_p = ext ? S0[(size_t)gj * nx + gi] : T(0);
    __syncthreads();
    // the faces: x face f (< R2FX) at ring row f / (RTX + 3), cell i0 - 2 + f % (RTX + 3);
    // y face g = f - R2FX at cell j0 - 2 + g / REX, ring column g % REX.  Past
    // the domain edge the LDS plane holds the boundary cell, as the ghosts do.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int f = t + h * RNT;
        if (f < R2F) {
            const bool fx = f < R2FX;
            const int g = fx ? f : f - R2FX;
            const int r = fx ? g / (RTX + 3) + 3 : g / REX + 2;   // LDS row of the face's cell m
            const int c = fx ? g % (RTX + 3) + 2 : g % REX + 3;   // and column
            const int st = fx ? 1 : RPX;
            const T *v = &sphi[r][c];
            const T vm2 = v[-2 * st], vm1 = v[-st], v0 = v[0], vp1 = v[st], vp2 = v[2 * st], vp3 = v[3 * st];
            T plus, minus;
            weno_hj_face(vm2 - T(2) * vm1 + v0, vm1 - T(2) * v0 + vp1, v0 - T(2) * vp1 + vp2,
                         vp1 - T(2) * vp2 + vp3, c3, plus, minus);
            fplus[f] = plus;
            fminus[f] = minus;
        }
    }
    __syncthreads();
    T grad_p = T(0), del_p = T(0), phi_p = T(0);
    // a ring cell's gradient only enters lambda through s2, s3 = 0 where delta = 0
    if (ext && !inner && !(BTN == 1 && in_band(sphi[pr][pc] * inv_w))) {
        if (BTN == 1) {
            ss2[ey][ex] = T(0);
            ss3[ey][ex] = T(0);
        }
    } else if (ext) {
        const T c0 = sphi[pr][pc];
        // a cell past the domain edge takes the boundary cell's values (gi, gj):
        // at ring position (cx, cy) its x faces i -/+ 1/2 are x faces
        // cy (RTX + 3) + cx (+ 1), its y faces j -/+ 1/2 y faces cy REX + cx (+ REX)
        const int cx = gi - i0 + 1, cy = gj - j0 + 1;
        const int fxm = cy * (RTX + 3) + cx, fym = R2FX + cy * REX + cx;
        const T ax = inv12dx * central12(sphi[pr][pc - 2], sphi[pr][pc - 1], sphi[pr][pc + 1], sphi[pr][pc + 2]);
        const T ay = inv12dy * central12(sphi[pr - 2][pc], sphi[pr - 1][pc], sphi[pr + 1][pc], sphi[pr + 2][pc]);
        const T xs1 = ax + inv12dx * fplus[fxm + 1], xs2 = ax - inv12dx * fminus[fxm];
        const T ys1 = ay + inv12dy * fplus[fym + REX], ys2 = ay - inv12dy * fminus[fym];
        grad_p = godunov<T, 2>(xs1, xs2, ys1, ys2, T(0), T(0), s0_p);
        phi_p = c0;
        if (BTN == 1) {
            T H;
            ls_funs_fast(c0, inv_w, H, del_p);
            ss2[ey][ex] = del_p * s0_p * (grad_p - T(1));
            ss3[ey][ex] = grad_p * del_p * del_p;
        }
    }
    if (BTN == 1) __syncthreads();
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
        if (t == 0 && m > 0.0) atomic_max_nonneg(ctl_slot(ctl, blockIdx.y * gridDim.x + blockIdx.x), m);
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

// Device-side control of the pseudo-time loop (level_set_rk3_redis.f90:17-23, 68):
// ctl[1] stop flag, ctl[2] pseudo-steps run, ctl[3] max |dphi| of the last
// one; the running pseudo-step's max |dphi| (bits of a non-negative double)
// is striped over ctl[4 .. 4 + CTL_SLOTS).  One wave: op 0 resets, op 2 closes
// a pseudo-step -- reduces the stripes to its max |dphi|, clears them for the
// next step and stops the loop once the step's max |dphi| <= 1e-6 (a NaN
// never stops it, as in the reference); the time limit is the host's count of
// launched steps.  Stage kernels return at once when the flag is set, so
// steps can be queued ahead without host round trips and the step count stays
// the reference's.
__global__ void k_redis_ctl(unsigned long long *ctl, int op)
{
    const int t = threadIdx.x;   // stripe t
    if (op == 0) {
        ctl[4 + t] = 0;
        if (t < 4) ctl[t] = 0;
        return;
    }
    if (ctl[1]) return;
    const double v = __longlong_as_double((long long)ctl[4 + t]);
    ctl[4 + t] = 0;
    // non-negative doubles (and a NaN, whose bits order above them) keep their
    // order as integers: the max of the bit patterns
    unsigned long long m = (unsigned long long)__double_as_longlong(v);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    if (t == 0) {
        ctl[2] += 1;
        ctl[3] = m;
        const double e = __longlong_as_double((long long)m);
        if (e <= 1.0e-6) ctl[1] = 1;
    }
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
// A, B, AA, BB, each (3, nb) column-major, then 1/Den of each row.
//
// Checkpointed line solve (ccd_solvers.f90:412-427): the reference's
// forward elimination and block substitution with its operation order per
// row.  One lane per line.  The forward
// elimination keeps its state (S, SS) in registers and stores it only at the
// start of every CSEG-row segment; the backward substitution walks the
// segments downwards, recomputes each segment's forward states from its
// checkpoint into registers and substitutes through them.  Each segment's
// loads are issued together, so memory latency is paid once per segment,
// and the per-row scratch round trip of a sweep that stores every row (32 bytes per cell in
// FP64) disappears.  The coefficient tables are staged in LDS when they fit
// (TLDS), else read from global memory (uniform, scalar loads).
//
// Line L starts at (L / lw) * ostride + (L % lw) * istride, element stride
// estride.  Checkpoint k of line L: ck[(2 k + comp) * NL + L], NL = lanes in
// the grid.
// ---------------------------------------------------------------------------
template <typename T, int DIM, bool FINAL, bool TLDS>
__global__ __launch_bounds__(256) void k_ccd_seg(const T *__restrict__ phi, const T *__restrict__ coef,
                                                  T *__restrict__ ck, T *__restrict__ out,
                                                  const T *__restrict__ dxv, const T *__restrict__ dyv,
                                                  T *__restrict__ gr, double *__restrict__ partial, int nlines,
                                                  int n, int lw, long ostride, long istride, long estride,
                                                  CcdArgs ca, const int *__restrict__ lines)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char ccd_lds[];
    const int nb = n + 6;
    const T *tab = coef;
    if (TLDS) {
        T *t = reinterpret_cast<T *>(ccd_lds);
        for (int i = threadIdx.x; i < 13 * nb; i += blockDim.x) t[i] = coef[i];
        __syncthreads();
        tab = t;
    }
    const T *A = tab, *B = tab + 3 * nb, *AA = tab + 6 * nb, *BB = tab + 9 * nb, *IDEN = tab + 12 * nb;
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = L < nlines;
    const int Lc = act ? L : nlines - 1;
    const int line = lines ? lines[Lc] : Lc;   // with a list, lane L solves listed line L of nlines
    const size_t base = (size_t)(line / lw) * (size_t)ostride + (size_t)(line % lw) * (size_t)istride;
    const size_t NL = (size_t)gridDim.x * blockDim.x;
    const T idx = T(1) / (T)ca.dx, idx2 = T(1) / ((T)ca.dx * (T)ca.dx);
    auto ldq = [&](int q) -> T { return phi[base + (size_t)clampi(q, 0, n - 1) * (size_t)estride]; };
    // forward row b from the window f(b-1), f(b), f(b+1) = phi(b-4), phi(b-3), phi(b-2)
    auto frow = [&](int b, T fm1, T f0, T fp1, T &S, T &SS) {
        T s = T(15.0 / 16.0) * (-fm1 + fp1) * idx;
        T ss = (T(3) * fm1 - T(6) * f0 + T(3) * fp1) * idx2;
        s = s - (A[3 * b] * S + B[3 * b] * SS);
        ss = ss - (AA[3 * b] * S + BB[3 * b] * SS);
        S = s;
        SS = ss;
    };
    double acc[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    auto emit = [&](int q, T d, T dxa, T dya, T pa) {
        const size_t a = base + (size_t)q * (size_t)estride;
        if (!FINAL) {
            out[a] = d;
        } else {
            T g;
            if (DIM == 3) g = sqrt(sq(dxa) + sq(dya) + sq(d));
            else g = sqrt(sq(dxa) + sq(d));
            gr[a] = g;
            // the mass integral is formed in FP64 in both forms
            double h, dd;
            ls_funs<double>((double)pa, ca.ls_wid, h, dd);
            const double gg = (double)g;
            const double r12 = ca.rho_12;
            double rho, plam;
            csum_add(acc[0], cmp[0], h * ca.dv);
            if (!ca.inverse) rho = h + r12 * (1.0 - h);
            else rho = r12 * h + (1.0 - h);
            csum_add(acc[1], cmp[1], rho * h * ca.dv);
            plam = dd * dd * gg;
            if (ca.inverse) plam = plam * (2.0 * (r12 - 1.0) * h + 1.0) * ca.dx * ca.dy * ca.dz;
            else plam = plam * (2.0 * (1.0 - r12) * h + r12) * ca.dx * ca.dy * ca.dz;
            csum_add(acc[2], cmp[2], plam);
        }
    };

    // row 0: one-sided closure (the state the reference's forward loop starts from)
    T S0, SS0;
    {
        const T a0 = ldq(-3), a1 = ldq(-2), a2 = ldq(-1), a3 = ldq(0);
        S0 = (T(-3.5) * a0 + T(4) * a1 - T(0.5) * a2) * idx;
        SS0 = (T(34.0 / 3.0) * a0 - T(83.0 / 4.0) * a1 + T(10) * a2 - T(7.0 / 12.0) * a3) * idx2;
    }
    // forward elimination over rows 1 .. nb-2, segment k = rows [k CSEG, k CSEG + CSEG - 1]
    const int K = (nb - 2) / CSEG + 1;   // segments holding rows 0 .. nb-2
    T S = S0, SS = SS0;
    for (int k = 0; k < K; ++k) {
        const int lo = k * CSEG;
        if (k > 0 && act) {   // state before segment k
            ck[(size_t)(2 * k) * NL + L] = S;
            ck[(size_t)(2 * k + 1) * NL + L] = SS;
        }
        T f[CSEG + 2];   // phi(b-4) for b = lo .. lo + CSEG + 1
#pragma unroll
        for (int r = 0; r < CSEG + 2; ++r) f[r] = ldq(lo - 4 + r);
#pragma unroll
        for (int r = 0; r < CSEG; ++r) {
            const int b = lo + r;
            if (b >= 1 && b <= nb - 2) frow(b, f[r], f[r + 1], f[r + 2], S, SS);
        }
    }
    // closure row nb-1 (reads box points nb-1 .. nb-4) and its 2x2 solve
    {
        const int b = nb - 1;
        const T e0 = ldq(b - 3), e1 = ldq(b - 4), e2 = ldq(b - 5), e3 = ldq(b - 6);
        T s = -(T(-3.5) * e0 + T(4) * e1 - T(0.5) * e2) * idx;
        T ss = (T(34.0 / 3.0) * e0 - T(83.0 / 4.0) * e1 + T(10) * e2 - T(7.0 / 12.0) * e3) * idx2;
        s = s - (A[3 * b] * S + B[3 * b] * SS);
        ss = ss - (AA[3 * b] * S + BB[3 * b] * SS);
        const T iden = IDEN[b];
        const T solS = -(B[3 * b + 1] * ss - BB[3 * b + 1] * s) * iden;
        const T solSS = (A[3 * b + 1] * ss - AA[3 * b + 1] * s) * iden;
        S = solS;
        SS = solSS;
    }
    // backward substitution, segment by segment from the top
    for (int k = K - 1; k >= 0; --k) {
        const int lo = k * CSEG, hi = min(lo + CSEG - 1, nb - 2);
        T f[CSEG + 2], dxa[CSEG], dya[CSEG], pa[CSEG];
#pragma unroll
        for (int r = 0; r < CSEG + 2; ++r) f[r] = ldq(lo - 4 + r);
        if (FINAL) {
#pragma unroll
            for (int r = 0; r < CSEG; ++r) {   // operands of the emitted point q = b - 3
                const size_t a = base + (size_t)clampi(lo + r - 3, 0, n - 1) * (size_t)estride;
                dxa[r] = dxv[a];
                dya[r] = DIM == 3 ? dyv[a] : T(0);
                pa[r] = phi[a];
            }
        }
        T fs[CSEG], fss[CSEG];   // forward states of the segment's rows
        T cs = S0, css = SS0;
        if (k > 0) {
            cs = ck[(size_t)(2 * k) * NL + Lc];
            css = ck[(size_t)(2 * k + 1) * NL + Lc];
        }
#pragma unroll
        for (int r = 0; r < CSEG; ++r) {
            const int b = lo + r;
            if (b >= 1 && b <= hi) frow(b, f[r], f[r + 1], f[r + 2], cs, css);
            fs[r] = cs;
            fss[r] = css;
        }
#pragma unroll
        for (int r = CSEG - 1; r >= 0; --r) {
            const int b = lo + r;
            if (b <= hi) {
                const T s = fs[r] - (A[3 * b + 2] * S + B[3 * b + 2] * SS);
                const T ss = fss[r] - (AA[3 * b + 2] * S + BB[3 * b + 2] * SS);
                const T iden = IDEN[b];
                const T solS = -(B[3 * b + 1] * ss - BB[3 * b + 1] * s) * iden;
                const T solSS = (A[3 * b + 1] * ss - AA[3 * b + 1] * s) * iden;
                S = solS;
                SS = solSS;
                if (act && b - 3 >= 0 && b - 3 < n) emit(b - 3, S, dxa[r], dya[r], pa[r]);
            }
        }
    }
    if (FINAL) {
        block_sum_comp<3>(acc, cmp);
        if (threadIdx.x == 0) {
            partial[3 * blockIdx.x + 0] = acc[0];
            partial[3 * blockIdx.x + 1] = acc[1];
            partial[3 * blockIdx.x + 2] = acc[2];
        }
    }
}

// CCD line solve for few lines (the 2D grids): LPL = 32 lanes per line, two
// lines per wave, or for lines of 541..1084 cells LPL = 64, one line per wave.
// With the factorisation fixed per row, both sweeps are affine
// recurrences whose 2x2 matrices are the same for every line: forward
// y_b = r_b - M_b y_{b-1}, M_b = [[A, B], [AA, BB]] of row b (the frow
// update); backward x_b = Q_b (f_b - N_b x_{b+1}), N_b the row's third
// block and Q_b its 2x2 solve (TWIN_BKS, matrix_solver.f90:47-78).  Lane s
// owns the forward rows lo..hi (at most KCC of rows 1..nb-2).  Each lane
// composes its rows into one affine map, an inclusive scan over the LPL lanes
// (five or six shuffle rounds) gives every lane the state entering its rows, and
// the lane re-runs its rows from that state with the reference's per-row
// arithmetic; the backward sweep does the same from the closure row down.
// Only the states entering each lane's rows carry the scan's rounding.
template <typename T, int DIM, bool FINAL, bool TLDS, int KC, int LPL>
__global__ __launch_bounds__(256) void k_ccd_part(const T *__restrict__ phi, const T *__restrict__ coef,
                                                   T *__restrict__ out, const T *__restrict__ dxv,
                                                   const T *__restrict__ dyv, T *__restrict__ gr,
                                                   double *__restrict__ partial, int nlines, int n, int lw,
                                                   long ostride, long istride, long estride, CcdArgs ca,
                                                   const int *__restrict__ lines)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char ccd_lds[];
    const int nb = n + 6;
    const T *tab = coef;
    if (TLDS) {
        T *t = reinterpret_cast<T *>(ccd_lds);
        for (int i = threadIdx.x; i < 13 * nb; i += blockDim.x) t[i] = coef[i];
        __syncthreads();
        tab = t;
    }
    const T *A = tab, *B = tab + 3 * nb, *AA = tab + 6 * nb, *BB = tab + 9 * nb, *IDEN = tab + 12 * nb;
    static_assert(LPL == 32 || LPL == 64, "a half or a whole wave per line");
    const int lane = threadIdx.x & 63, s = lane & (LPL - 1), half = lane & (64 - LPL);
    const int line = (int)((xcd_block() * blockDim.x + threadIdx.x) >> 6) * (64 / LPL) + (LPL == 32 ? lane >> 5 : 0);
    const bool act = line < nlines;
    const int Lc = act ? line : nlines - 1;   // idle halves shadow the last line and emit nothing
    const int bl = lines ? lines[Lc] : Lc;     // with a list, half `line` solves listed line `line` of nlines
    const size_t base = (size_t)(bl / lw) * (size_t)ostride + (size_t)(bl % lw) * (size_t)istride;
    const T idx = T(1) / (T)ca.dx, idx2 = T(1) / ((T)ca.dx * (T)ca.dx);
    auto ldq = [&](int q) -> T { return phi[base + (size_t)clampi(q, 0, n - 1) * (size_t)estride]; };
    const int NR = nb - 2, K = (NR + LPL - 1) / LPL;
    const int lo = 1 + s * K, hi = min(lo + K - 1, NR);
    const int nrow = hi >= lo ? hi - lo + 1 : 0;

    // row 0: one-sided closure, the state the forward sweep starts from
    T S0, SS0;
    {
        const T a0 = ldq(-3), a1 = ldq(-2), a2 = ldq(-1), a3 = ldq(0);
        S0 = (T(-3.5) * a0 + T(4) * a1 - T(0.5) * a2) * idx;
        SS0 = (T(34.0 / 3.0) * a0 - T(83.0 / 4.0) * a1 + T(10) * a2 - T(7.0 / 12.0) * a3) * idx2;
    }
    T f[KC + 2];   // phi(b-4) for b = lo .. lo + KC + 1
#pragma unroll
    for (int r = 0; r < KC + 2; ++r) f[r] = ldq(lo - 4 + r);
    auto rhs = [&](int r, T &rs, T &rss) {   // the row's right-hand side from f(b-1), f(b), f(b+1)
        rs = T(15.0 / 16.0) * (-f[r] + f[r + 2]) * idx;
        rss = (T(3) * f[r] - T(6) * f[r + 1] + T(3) * f[r + 2]) * idx2;
    };

    // forward: this lane's rows as y_out = P y_in + v
    T p00 = T(1), p01 = T(0), p10 = T(0), p11 = T(1), v0 = T(0), v1 = T(0);
#pragma unroll
    for (int r = 0; r < KC; ++r) {
        if (r < nrow) {
            const int b = lo + r;
            const T m00 = A[3 * b], m01 = B[3 * b], m10 = AA[3 * b], m11 = BB[3 * b];
            T rs, rss;
            rhs(r, rs, rss);
            const T nv0 = rs - (m00 * v0 + m01 * v1), nv1 = rss - (m10 * v0 + m11 * v1);
            const T q00 = -(m00 * p00 + m01 * p10), q01 = -(m00 * p01 + m01 * p11);
            const T q10 = -(m10 * p00 + m11 * p10), q11 = -(m10 * p01 + m11 * p11);
            v0 = nv0; v1 = nv1;
            p00 = q00; p01 = q01; p10 = q10; p11 = q11;
        }
    }
    // inclusive scan: map of lanes 0..s composed
#pragma unroll
    for (int d = 1; d < LPL; d <<= 1) {
        const int src = s >= d ? lane - d : lane;
        const T a00 = __shfl(p00, src, 64), a01 = __shfl(p01, src, 64), a10 = __shfl(p10, src, 64);
        const T a11 = __shfl(p11, src, 64), w0 = __shfl(v0, src, 64), w1 = __shfl(v1, src, 64);
        if (s >= d) {   // this ∘ earlier
            const T nv0 = p00 * w0 + p01 * w1 + v0, nv1 = p10 * w0 + p11 * w1 + v1;
            const T q00 = p00 * a00 + p01 * a10, q01 = p00 * a01 + p01 * a11;
            const T q10 = p10 * a00 + p11 * a10, q11 = p10 * a01 + p11 * a11;
            v0 = nv0; v1 = nv1;
            p00 = q00; p01 = q01; p10 = q10; p11 = q11;
        }
    }
    const T e0 = p00 * S0 + p01 * SS0 + v0, e1 = p10 * S0 + p11 * SS0 + v1;   // state after this lane's rows
    const T in0 = __shfl(e0, (lane - 1) & 63, 64), in1 = __shfl(e1, (lane - 1) & 63, 64);
    const T last0 = __shfl(e0, half + LPL - 1, 64), last1 = __shfl(e1, half + LPL - 1, 64);   // after row nb-2
    // this lane's forward states with the reference's per-row arithmetic
    T fs[KC], fss[KC];
    {
        T cs = s == 0 ? S0 : in0, css = s == 0 ? SS0 : in1;
#pragma unroll
        for (int r = 0; r < KC; ++r) {
            if (r < nrow) {
                const int b = lo + r;
                T rs, rss;
                rhs(r, rs, rss);
                rs = rs - (A[3 * b] * cs + B[3 * b] * css);
                rss = rss - (AA[3 * b] * cs + BB[3 * b] * css);
                cs = rs;
                css = rss;
            }
            fs[r] = cs;
            fss[r] = css;
        }
    }
    // closure row nb-1 and its 2x2 solve (the same on every lane of the half)
    T XS, XSS;
    {
        const int b = nb - 1;
        const T g0 = ldq(b - 3), g1 = ldq(b - 4), g2 = ldq(b - 5), g3 = ldq(b - 6);
        T s_ = -(T(-3.5) * g0 + T(4) * g1 - T(0.5) * g2) * idx;
        T ss_ = (T(34.0 / 3.0) * g0 - T(83.0 / 4.0) * g1 + T(10) * g2 - T(7.0 / 12.0) * g3) * idx2;
        s_ = s_ - (A[3 * b] * last0 + B[3 * b] * last1);
        ss_ = ss_ - (AA[3 * b] * last0 + BB[3 * b] * last1);
        const T iden = IDEN[b];
        XS = -(B[3 * b + 1] * ss_ - BB[3 * b + 1] * s_) * iden;
        XSS = (A[3 * b + 1] * ss_ - AA[3 * b + 1] * s_) * iden;
    }
    // backward: this lane's rows as x_lo = G x_{hi+1} + w
    T g00 = T(1), g01 = T(0), g10 = T(0), g11 = T(1), w0 = T(0), w1 = T(0);
#pragma unroll
    for (int r = KC - 1; r >= 0; --r) {
        if (r < nrow) {
            const int b = lo + r;
            const T n00 = A[3 * b + 2], n01 = B[3 * b + 2], n10 = AA[3 * b + 2], n11 = BB[3 * b + 2];
            const T iden = IDEN[b];
            const T q00 = BB[3 * b + 1] * iden, q01 = -B[3 * b + 1] * iden;
            const T q10 = -AA[3 * b + 1] * iden, q11 = A[3 * b + 1] * iden;
            // w <- Q (f - N w), G <- -Q N G
            const T t0 = fs[r] - (n00 * w0 + n01 * w1), t1 = fss[r] - (n10 * w0 + n11 * w1);
            const T nw0 = q00 * t0 + q01 * t1, nw1 = q10 * t0 + q11 * t1;
            const T k00 = n00 * g00 + n01 * g10, k01 = n00 * g01 + n01 * g11;
            const T k10 = n10 * g00 + n11 * g10, k11 = n10 * g01 + n11 * g11;
            g00 = -(q00 * k00 + q01 * k10); g01 = -(q00 * k01 + q01 * k11);
            g10 = -(q10 * k00 + q11 * k10); g11 = -(q10 * k01 + q11 * k11);
            w0 = nw0; w1 = nw1;
        }
    }
    // inclusive suffix scan: map of lanes s..LPL-1 composed (this ∘ later)
#pragma unroll
    for (int d = 1; d < LPL; d <<= 1) {
        const int src = s + d < LPL ? lane + d : lane;
        const T a00 = __shfl(g00, src, 64), a01 = __shfl(g01, src, 64), a10 = __shfl(g10, src, 64);
        const T a11 = __shfl(g11, src, 64), u0 = __shfl(w0, src, 64), u1 = __shfl(w1, src, 64);
        if (s + d < LPL) {
            const T nw0 = g00 * u0 + g01 * u1 + w0, nw1 = g10 * u0 + g11 * u1 + w1;
            const T q00 = g00 * a00 + g01 * a10, q01 = g00 * a01 + g01 * a11;
            const T q10 = g10 * a00 + g11 * a10, q11 = g10 * a01 + g11 * a11;
            w0 = nw0; w1 = nw1;
            g00 = q00; g01 = q01; g10 = q10; g11 = q11;
        }
    }
    const T lo0 = g00 * XS + g01 * XSS + w0, lo1 = g10 * XS + g11 * XSS + w1;   // x at this lane's first row
    const T nx0 = __shfl(lo0, (lane + 1) & 63, 64), nx1 = __shfl(lo1, (lane + 1) & 63, 64);
    // this lane's rows, top down, with the reference's arithmetic from x_{hi+1}
    double acc[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    T S = s == LPL - 1 ? XS : nx0, SS = s == LPL - 1 ? XSS : nx1;
#pragma unroll
    for (int r = KC - 1; r >= 0; --r) {
        if (r < nrow) {
            const int b = lo + r;
            const T s_ = fs[r] - (A[3 * b + 2] * S + B[3 * b + 2] * SS);
            const T ss_ = fss[r] - (AA[3 * b + 2] * S + BB[3 * b + 2] * SS);
            const T iden = IDEN[b];
            const T solS = -(B[3 * b + 1] * ss_ - BB[3 * b + 1] * s_) * iden;
            const T solSS = (A[3 * b + 1] * ss_ - AA[3 * b + 1] * s_) * iden;
            S = solS;
            SS = solSS;
            const int q = b - 3;
            if (act && q >= 0 && q < n) {
                const size_t a = base + (size_t)q * (size_t)estride;
                if (!FINAL) {
                    out[a] = S;
                } else {
                    const T dxa = dxv[a], dya = DIM == 3 ? dyv[a] : T(0), pa = phi[a];
                    T g;
                    if (DIM == 3) g = sqrt(sq(dxa) + sq(dya) + sq(S));
                    else g = sqrt(sq(dxa) + sq(S));
                    gr[a] = g;
                    double h, dd;
                    ls_funs<double>((double)pa, ca.ls_wid, h, dd);
                    const double gg = (double)g;
                    const double r12 = ca.rho_12;
                    double rho, plam;
                    csum_add(acc[0], cmp[0], h * ca.dv);
                    if (!ca.inverse) rho = h + r12 * (1.0 - h);
                    else rho = r12 * h + (1.0 - h);
                    csum_add(acc[1], cmp[1], rho * h * ca.dv);
                    plam = dd * dd * gg;
                    if (ca.inverse) plam = plam * (2.0 * (r12 - 1.0) * h + 1.0) * ca.dx * ca.dy * ca.dz;
                    else plam = plam * (2.0 * (1.0 - r12) * h + r12) * ca.dx * ca.dy * ca.dz;
                    csum_add(acc[2], cmp[2], plam);
                }
            }
        }
    }
    if (FINAL) {
        block_sum_comp<3>(acc, cmp);
        if (threadIdx.x == 0) {
            partial[3 * blockIdx.x + 0] = acc[0];
            partial[3 * blockIdx.x + 1] = acc[1];
            partial[3 * blockIdx.x + 2] = acc[2];
        }
    }
}

// ---------------------------------------------------------------------------
// UCCD advection (level_set_rk3_source_ccd, level_set_rk3_solver.f90:223-286;
// root_solve "uccd", ccd_solvers.f90:428-438): per line, the upwind (U) and
// downwind (D) block systems of asgn_src_uccd (:312-345) with their TWIN_DEC
// tables (coef: U tables, then D tables, 13 nb each), solved like
// k_ccd_seg (checkpoints of both states every CSEG rows: ck[(4 k + comp) * NL
// + L]); the emitted derivative at an interior point is U's for u > 0 and D's
// otherwise.  The RHS accumulates -u phi_x (MODE 0), - v phi_y (MODE 1) and,
// in the last direction (MODE 2), - w phi_z followed by the RK stage of
// level_set_rk3_solver.f90:16-88, written to phi_out: the segment recompute
// re-reads phi, so the stage cannot update it in place.
// ---------------------------------------------------------------------------
template <typename T, int MODE, int STAGE, bool TLDS>
__global__ __launch_bounds__(256) void k_uccd_seg(const T *__restrict__ phi, T *__restrict__ phi_out,
                                                   const T *__restrict__ vel, T *R, T *L1, T *L2,
                                                   const T *__restrict__ coef, T *__restrict__ ck, int nlines,
                                                   int n, int lw, long ostride, long istride, long estride, T dx,
                                                   T dt)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char uccd_lds[];
    const int nb = n + 6;
    const T *tab = coef;
    if (TLDS) {
        T *t = reinterpret_cast<T *>(uccd_lds);
        for (int i = threadIdx.x; i < 26 * nb; i += blockDim.x) t[i] = coef[i];
        __syncthreads();
        tab = t;
    }
    const T *AU = tab, *BU = tab + 3 * nb, *AAU = tab + 6 * nb, *BBU = tab + 9 * nb, *IU = tab + 12 * nb;
    const T *AD = tab + 13 * nb, *BD = AD + 3 * nb, *AAD = AD + 6 * nb, *BBD = AD + 9 * nb, *ID = AD + 12 * nb;
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = L < nlines;
    const int Lc = act ? L : nlines - 1;
    const size_t base = (size_t)(Lc / lw) * (size_t)ostride + (size_t)(Lc % lw) * (size_t)istride;
    const size_t NL = (size_t)gridDim.x * blockDim.x;
    const T idx = T(1) / dx, idx2 = T(1) / (dx * dx);
    const T c3 = T(-0.06096119008109), c2 = T(1.99692238016218), c1 = T(-1.93596119008109);
    auto ldq = [&](int q) -> T { return phi[base + (size_t)clampi(q, 0, n - 1) * (size_t)estride]; };
    // forward row b of both systems from f(b-1), f(b), f(b+1) = phi(b-4), phi(b-3), phi(b-2)
    auto frow = [&](int b, T fm1, T f0, T fp1, T &SU, T &SSU, T &SD, T &SSD) {
        const T ss0 = (T(3) * fm1 - T(6) * f0 + T(3) * fp1) * idx2;
        T su = (c1 * fm1 + c2 * f0 + c3 * fp1) * idx;
        T sd = -(c3 * fm1 + c2 * f0 + c1 * fp1) * idx;
        su = su - (AU[3 * b] * SU + BU[3 * b] * SSU);
        const T ssu = ss0 - (AAU[3 * b] * SU + BBU[3 * b] * SSU);
        sd = sd - (AD[3 * b] * SD + BD[3 * b] * SSD);
        const T ssd = ss0 - (AAD[3 * b] * SD + BBD[3 * b] * SSD);
        SU = su; SSU = ssu; SD = sd; SSD = ssd;
    };
    auto emit = [&](int q, T d1, T uq, T Rv, T Pv, T L1v, T L2v) {
        const size_t a = base + (size_t)q * (size_t)estride;
        if (MODE == 0) {
            R[a] = -uq * d1;
        } else if (MODE == 1) {
            R[a] = Rv - uq * d1;
        } else {
            const T r = Rv - uq * d1;
            if (STAGE == 1) {
                L1[a] = r;
                phi_out[a] = Pv + r * dt;
            } else if (STAGE == 2) {
                L2[a] = r;
                phi_out[a] = Pv + (T(-3) * L1v + r) / T(4) * dt;
            } else {
                phi_out[a] = Pv + (-L1v - L2v + T(8) * r) / T(12) * dt;
            }
        }
    };

    // row 0: the one-sided closure, the same for both systems
    T S0, SS0;
    {
        const T a0 = ldq(-3), a1 = ldq(-2), a2 = ldq(-1), a3 = ldq(0);
        S0 = (T(-3.5) * a0 + T(4) * a1 - T(0.5) * a2) * idx;
        SS0 = (T(34.0 / 3.0) * a0 - T(83.0 / 4.0) * a1 + T(10) * a2 - T(7.0 / 12.0) * a3) * idx2;
    }
    const int K = (nb - 2) / CSEG + 1;
    T SU = S0, SSU = SS0, SD = S0, SSD = SS0;
    for (int k = 0; k < K; ++k) {
        const int lo = k * CSEG;
        if (k > 0 && act) {
            ck[(size_t)(4 * k) * NL + L] = SU;
            ck[(size_t)(4 * k + 1) * NL + L] = SSU;
            ck[(size_t)(4 * k + 2) * NL + L] = SD;
            ck[(size_t)(4 * k + 3) * NL + L] = SSD;
        }
        T f[CSEG + 2];
#pragma unroll
        for (int r = 0; r < CSEG + 2; ++r) f[r] = ldq(lo - 4 + r);
#pragma unroll
        for (int r = 0; r < CSEG; ++r) {
            const int b = lo + r;
            if (b >= 1 && b <= nb - 2) frow(b, f[r], f[r + 1], f[r + 2], SU, SSU, SD, SSD);
        }
    }
    // closure row nb-1 of both systems and their 2x2 solves
    {
        const int b = nb - 1;
        const T e0 = ldq(b - 3), e1 = ldq(b - 4), e2 = ldq(b - 5), e3 = ldq(b - 6);
        const T s0 = -(T(-3.5) * e0 + T(4) * e1 - T(0.5) * e2) * idx;
        const T ss0 = (T(34.0 / 3.0) * e0 - T(83.0 / 4.0) * e1 + T(10) * e2 - T(7.0 / 12.0) * e3) * idx2;
        {
            const T s = s0 - (AU[3 * b] * SU + BU[3 * b] * SSU);
            const T ss = ss0 - (AAU[3 * b] * SU + BBU[3 * b] * SSU);
            SU = -(BU[3 * b + 1] * ss - BBU[3 * b + 1] * s) * IU[b];
            SSU = (AU[3 * b + 1] * ss - AAU[3 * b + 1] * s) * IU[b];
        }
        {
            const T s = s0 - (AD[3 * b] * SD + BD[3 * b] * SSD);
            const T ss = ss0 - (AAD[3 * b] * SD + BBD[3 * b] * SSD);
            SD = -(BD[3 * b + 1] * ss - BBD[3 * b + 1] * s) * ID[b];
            SSD = (AD[3 * b + 1] * ss - AAD[3 * b + 1] * s) * ID[b];
        }
    }
    for (int k = K - 1; k >= 0; --k) {
        const int lo = k * CSEG, hi = min(lo + CSEG - 1, nb - 2);
        T f[CSEG + 2], uq[CSEG], Rq[CSEG], Pq[CSEG], L1q[CSEG], L2q[CSEG];
#pragma unroll
        for (int r = 0; r < CSEG + 2; ++r) f[r] = ldq(lo - 4 + r);
#pragma unroll
        for (int r = 0; r < CSEG; ++r) {   // operands of the emitted point q = b - 3
            const size_t a = base + (size_t)clampi(lo + r - 3, 0, n - 1) * (size_t)estride;
            uq[r] = vel[a];
            Rq[r] = MODE >= 1 ? R[a] : T(0);
            Pq[r] = MODE == 2 ? phi[a] : T(0);
            L1q[r] = (MODE == 2 && STAGE >= 2) ? L1[a] : T(0);
            L2q[r] = (MODE == 2 && STAGE == 3) ? L2[a] : T(0);
        }
        T fu[CSEG], fuu[CSEG], fd[CSEG], fdd[CSEG];
        T cu = S0, cuu = SS0, cd = S0, cdd = SS0;
        if (k > 0) {
            cu = ck[(size_t)(4 * k) * NL + Lc];
            cuu = ck[(size_t)(4 * k + 1) * NL + Lc];
            cd = ck[(size_t)(4 * k + 2) * NL + Lc];
            cdd = ck[(size_t)(4 * k + 3) * NL + Lc];
        }
#pragma unroll
        for (int r = 0; r < CSEG; ++r) {
            const int b = lo + r;
            if (b >= 1 && b <= hi) frow(b, f[r], f[r + 1], f[r + 2], cu, cuu, cd, cdd);
            fu[r] = cu; fuu[r] = cuu; fd[r] = cd; fdd[r] = cdd;
        }
#pragma unroll
        for (int r = CSEG - 1; r >= 0; --r) {
            const int b = lo + r;
            if (b <= hi) {
                {
                    const T s = fu[r] - (AU[3 * b + 2] * SU + BU[3 * b + 2] * SSU);
                    const T ss = fuu[r] - (AAU[3 * b + 2] * SU + BBU[3 * b + 2] * SSU);
                    const T sol = -(BU[3 * b + 1] * ss - BBU[3 * b + 1] * s) * IU[b];
                    SSU = (AU[3 * b + 1] * ss - AAU[3 * b + 1] * s) * IU[b];
                    SU = sol;
                }
                {
                    const T s = fd[r] - (AD[3 * b + 2] * SD + BD[3 * b + 2] * SSD);
                    const T ss = fdd[r] - (AAD[3 * b + 2] * SD + BBD[3 * b + 2] * SSD);
                    const T sol = -(BD[3 * b + 1] * ss - BBD[3 * b + 1] * s) * ID[b];
                    SSD = (AD[3 * b + 1] * ss - AAD[3 * b + 1] * s) * ID[b];
                    SD = sol;
                }
                if (act && b - 3 >= 0 && b - 3 < n)
                    emit(b - 3, uq[r] > T(0) ? SU : SD, uq[r], Rq[r], Pq[r], L1q[r], L2q[r]);
            }
        }
    }
}

// Deterministic final reduction of per-block partials (fixed order), then the
// mass-correction multiplier lambda = (imass - mass) / sum (mass_preserving_level_set.f90:46).
// scal: [0] vol, [1] mass, [2] denominator, [3] lambda, [4] flag (1 = empty band).
__global__ void k_finalize3(const double *__restrict__ partial, int nblk, double imass, double *scal)
{
    double acc[3];
    sum_partials3(partial, nblk, acc);
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
    const T p = phi[q];
    double H, D;
    ls_funs<double>((double)p, (double)ls_wid, H, D);
    if (D == 0.0) return;   // phi + lambda 0 |grad phi| is phi: no read of |grad phi|, no store
    phi[q] = (T)((double)p + lam * D * (double)gr[q]);
}

// The same update on the cells of the listed lines (every cell where delta can
// be nonzero lies on one); thread t updates cell t / nlist of listed line
// t % nlist, so neighbouring lanes take neighbouring lines.
template <typename T>
__global__ void k_mass_update_lines(T *__restrict__ phi, const T *__restrict__ gr, T ls_wid,
                                    const double *__restrict__ scal, const int *__restrict__ lines, int nlist, int n,
                                    int lw, long ostride, long istride, long estride)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nlist * n) return;
    const int l = lines[t % (size_t)nlist], q = (int)(t / (size_t)nlist);
    const size_t a = (size_t)(l / lw) * (size_t)ostride + (size_t)(l % lw) * (size_t)istride +
                     (size_t)q * (size_t)estride;
    const double lam = scal[3];
    const T p = phi[a];
    double H, D;
    ls_funs<double>((double)p, (double)ls_wid, H, D);
    if (D == 0.0) return;
    phi[a] = (T)((double)p + lam * D * (double)gr[a]);
}

// Lines that cross the interface band, for the mass correction's CCD passes:
// flag x line (j, k), y line (i, k) and z line (i, j) of every cell with
// |phi| < wband.  wband = ls_wid (1 + 1e-6) covers every cell where ls_funs
// gives delta != 0 (|phi| <= ls_wid (1 - 1e-12)).  The flags must be zero on
// entry; fz is unused in 2D.
//
// TUBE (the narrow band, 3D): phi is first clamped to [-wband, wband], so the
// cells outside the tube hold a constant, which advection leaves as it is
// (with a frozen distance field there, the mismatch with the advected tube
// would grow into a jump at the tube's edge that the implicit line solves
// carry back towards the interface); H, delta and the sign do not change,
// wband being far beyond ls_wid.  The lines of the six neighbours of every
// band cell are flagged too: a clamped cell next to the tube is then updated
// with its full right-hand side, and the tube can grow into the clamped field.
template <typename T, bool TUBE>
__global__ void k_band_line_flags(T *__restrict__ phi, int nx, int ny, int nz, double wband,
                                  unsigned char *__restrict__ fx, unsigned char *__restrict__ fy,
                                  unsigned char *__restrict__ fz)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)nx * ny * nz) return;
    const double p = (double)phi[q];
    if (TUBE && fabs(p) > wband) phi[q] = (T)(p > 0.0 ? wband : -wband);
    if (!(fabs(p) < wband)) return;
    const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
    fx[j + (size_t)ny * k] = 1;
    fy[i + (size_t)nx * k] = 1;
    if (nz > 1) fz[i + (size_t)nx * j] = 1;
    if (TUBE) {
        // x line (j, k), y line (i, k), z line (i, j) of the neighbours
        const int im = max(i - 1, 0), ip = min(i + 1, nx - 1), jm = max(j - 1, 0), jp = min(j + 1, ny - 1);
        const int km = max(k - 1, 0), kp = min(k + 1, nz - 1);
        fx[jm + (size_t)ny * k] = 1;
        fx[jp + (size_t)ny * k] = 1;
        fx[j + (size_t)ny * km] = 1;
        fx[j + (size_t)ny * kp] = 1;
        fy[im + (size_t)nx * k] = 1;
        fy[ip + (size_t)nx * k] = 1;
        fy[i + (size_t)nx * km] = 1;
        fy[i + (size_t)nx * kp] = 1;
        fz[im + (size_t)nx * j] = 1;
        fz[ip + (size_t)nx * j] = 1;
        fz[i + (size_t)nx * jm] = 1;
        fz[i + (size_t)nx * jp] = 1;
    }
}

// Ordered lists of the flagged lines, one workgroup per direction d (blockIdx.x):
// each thread counts the flags of its run of 16-byte words, an exclusive scan
// over the workgroup gives its first slot, and it writes its lines in
// ascending order; count[d] receives the total.  The flags of direction d
// start at byte band_flag_offset(d) and their padding bytes are zero.
__global__ __launch_bounds__(1024) void k_compact_lines(const unsigned char *__restrict__ flags, int nfx, int nfy,
                                                        int nfz, int *__restrict__ lx, int *__restrict__ ly,
                                                        int *__restrict__ lz, unsigned *__restrict__ count)
{
    __shared__ int wsum[16];
    const int d = blockIdx.x;
    const int n = d == 0 ? nfx : d == 1 ? nfy : nfz;
    int *list = d == 0 ? lx : d == 1 ? ly : lz;
    const uint4 *f = reinterpret_cast<const uint4 *>(flags + band_flag_offset(d, nfx, nfy));
    const int nw = (n + 15) / 16, cw = (nw + 1023) / 1024;
    const int w0 = min((int)threadIdx.x * cw, nw), w1 = min(w0 + cw, nw);
    int cnt = 0;
    for (int w = w0; w < w1; ++w) {   // flags are 0 or 1: one set bit per flagged byte
        const uint4 v = f[w];
        cnt += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = cnt;
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
        const int y = __shfl_up(inc, k, 64);
        if (lane >= k) inc += y;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int wbase = 0;
    for (int k = 0; k < wid; ++k) wbase += wsum[k];
    int o = wbase + inc - cnt;
    for (int w = w0; w < w1; ++w) {
        const uint4 v = f[w];
        const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if ((u[e] >> (8 * b)) & 0xffu) list[o++] = w * 16 + e * 4 + b;
    }
    if (threadIdx.x == 1023) count[d] = (unsigned)(wbase + inc);
}

// Mass and volume sums (the ls_mv terms of the final CCD pass) of the lines
// without a band cell (lflag 0): delta = 0 on them, so they add nothing to the
// correction's denominator and need no solve.  One lane per line and segment
// of seg cells (blockIdx.y the segment): a lane walking a whole line keeps one
// wave per SIMD busy with a chain of dependent load batches.
template <typename T>
__global__ __launch_bounds__(256) void k_line_hsums(const T *__restrict__ phi, const unsigned char *__restrict__ lflag,
                                                    int nlines, int n, int lw, long ostride, long istride,
                                                    long estride, int seg, CcdArgs ca, double *__restrict__ partial)
{
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    const int qb = (int)blockIdx.y * seg, qe = min(qb + seg, n);
    double acc[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    if (L < nlines && !lflag[L]) {
        const size_t base = (size_t)(L / lw) * (size_t)ostride + (size_t)(L % lw) * (size_t)istride;
        const double r12 = ca.rho_12;
        for (int q0 = qb; q0 < qe; q0 += 16) {
            T v[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = phi[base + (size_t)min(q0 + r, qe - 1) * (size_t)estride];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (q0 + r < qe) {
                    double h, dd;
                    ls_funs<double>((double)v[r], ca.ls_wid, h, dd);
                    csum_add(acc[0], cmp[0], h * ca.dv);
                    const double rho = ca.inverse ? r12 * h + (1.0 - h) : h + r12 * (1.0 - h);
                    csum_add(acc[1], cmp[1], rho * h * ca.dv);
                }
            }
        }
    }
    block_sum_comp<3>(acc, cmp);
    if (threadIdx.x == 0) {
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        partial[3 * b + 0] = acc[0];
        partial[3 * b + 1] = acc[1];
        partial[3 * b + 2] = acc[2];
    }
}

__global__ void k_zero_bytes(unsigned char *__restrict__ p, size_t n)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

// LS mass and volume (ls_mv, ns_solver/others.f90:185-212) as block partials.
template <typename T>
__global__ void k_mass_vol(const T *__restrict__ phi, size_t n, CcdArgs ca, double *__restrict__ partial)
{
    double acc[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double h, D;
        ls_funs<double>((double)phi[q], ca.ls_wid, h, D);
        const double r12 = ca.rho_12;
        double rho;
        csum_add(acc[0], cmp[0], h * ca.dv);
        if (!ca.inverse) rho = h + r12 * (1.0 - h);
        else rho = r12 * h + (1.0 - h);
        csum_add(acc[1], cmp[1], rho * h * ca.dv);
    }
    block_sum_comp<3>(acc, cmp);
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
    double acc[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double src = heavyside<double>((double)phi0[q] / ls_wid);
        double tgt = heavyside<double>((double)phi[q] / ls_wid);
        csum_add(acc[0], cmp[0], fabs(src - tgt));
    }
    block_sum_comp<3>(acc, cmp);
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
__global__ __launch_bounds__(256) void k_velocity(T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
                                                  VelArgs va, const double *__restrict__ tab)
{
    // the (x, y) threads of a workgroup cover a block of rows of planes
    // k = blockIdx.z, blockIdx.z + gridDim.z, ...
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= va.nx) return;
    const double *tx = tab, *ty = tab + 2 * va.nx, *tz = ty + 2 * va.ny;
    for (int k = blockIdx.z; k < va.nz; k += gridDim.z)
        for (int j = blockIdx.y * blockDim.y + threadIdx.y; j < va.ny; j += gridDim.y * blockDim.y) {
            const size_t q = ((size_t)k * va.ny + j) * va.nx + i;
            double uu, vv, ww = 0.0;
            if (va.field == 0) {
                const double x = va.xstart + (double)((float)(i + 1) - 0.5f) * va.dx;
                const double y = va.ystart + (double)((float)(j + 1) - 0.5f) * va.dy;
                uu = -va.omega * (y - 0.5);
                vv = va.omega * (x - 0.5);
            } else {
                // sin(pi c) and sin(2 pi c) of each axis from the tables of k_vel_tables
                const double sx = tx[2 * i], s2x = tx[2 * i + 1], sy = ty[2 * j], s2y = ty[2 * j + 1];
                if (va.field == 1) {
                    uu = -sq(sx) * s2y * va.ct;
                    vv = sq(sy) * s2x * va.ct;
                } else {
                    const double sz = tz[2 * k], s2z = tz[2 * k + 1];
                    uu = 2.0 * sq(sx) * s2y * s2z * va.ct;
                    vv = -s2x * sq(sy) * s2z * va.ct;
                    ww = -s2x * s2y * sq(sz) * va.ct;
                }
            }
            u[q] = (T)uu;
            v[q] = (T)vv;
            w[q] = (T)ww;
        }
}

// sin(pi c) and sin(2 pi c) at the cell centres c of each axis
// (velocity.f90 evaluates them per cell; the products are formed in the same order).
__global__ void k_vel_tables(VelArgs va, double *__restrict__ tab)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = va.nx + va.ny + (va.field == 2 ? va.nz : 0);
    if (t >= nt) return;
    const double pi = 3.141592653589793238462643383279502884;
    double c;
    if (t < va.nx) c = va.xstart + (double)((float)(t + 1) - 0.5f) * va.dx;
    else if (t < va.nx + va.ny) c = va.ystart + (double)((float)(t - va.nx + 1) - 0.5f) * va.dy;
    else c = va.zstart + (double)((float)(t - va.nx - va.ny + 1) - 0.5f) * va.dz;
    tab[2 * t] = sin(pi * c);
    tab[2 * t + 1] = sin(2.0 * pi * c);
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
                     T *scr, int nx, int ny, int nz, T dd, T dt, const int *lines, int nlist, const TubeMask &tm)
{
    const int n = dir == 0 ? nx : dir == 1 ? ny : nz;
    const int nall = dir == 0 ? ny * nz : dir == 1 ? nx * nz : nx * ny;
    // line L starts at (L / lw) * ostride + (L % lw) * istride
    const int lw = dir == 0 ? nall : nx;
    const long ostride = dir == 0 ? 0 : dir == 1 ? (long)nx * ny : (long)nx;
    const long istride = dir == 0 ? (long)nx : 1;
    const long estride = dir == 0 ? 1 : dir == 1 ? (long)nx : (long)nx * ny;
    // Few lines: the paired and twisted kernels leave most SIMDs idle and each
    // lane walks the whole line, so the partitioned kernel (one wave per
    // line, rows in registers) takes over.  Measured: 2D 256^2 x pass
    // 129 -> 15 us, y pass 99 -> 21 us; 128^3 x pass 70 -> 67 us (FP32
    // 58 -> 49); but strided y and z lines of 128^3 (16384 lines) 50 -> 103 us,
    // its loads being one line per lane, and at 256^3 the paired x kernel is
    // as fast (FP64) or faster (FP32 182 against 220 us).  Listed lines (the
    // narrow band) take it up to PART_MAX_LISTED_ADV strided lines.
    int nlines = lines ? nlist : nall;
    // many lines: the tiled solve reads and writes each field once, coalesced;
    // lines of 514..1025 cells (2D 1024^2) take its long-line form whatever
    // their count, since the paired and twisted kernels would give them one
    // to four lanes each
    if ((n >= 34 && n <= 32 * PKX + 1 && nlines >= TILE_MIN_LINES) ||
        (n > 32 * PKX + 1 && n <= 64 * 16 + 1 && !tm.fx)) {
        launch_adv_tile<T>(st, mode, stage, phi, vel, R, L1, L2, nlines, n, lw, ostride, istride, estride, dd, dt,
                           lines, tm);
        return;
    }
    const bool part = n >= 34 && n <= 32 * PKX + 1 &&
                      nlines <= (dir == 0 ? PART_MAX_LINES_X : lines ? PART_MAX_LISTED_ADV : PART_MAX_LINES_STRIDED);
    if (mode == 0 && !part && lines) {   // the paired x kernel takes every line
        lines = nullptr;
        nlines = nall;
    }
    if (nlines == 0) return;
    const size_t threads = (size_t)((nlines + 31) / 32) * 64;       // two lanes per line (x)
    const size_t threads4 = (size_t)((nlines + 15) / 16) * 64;      // four lanes per line (y, z)
    const dim3 g((unsigned)((threads + 255) / 256)), g4((unsigned)((threads4 + 255) / 256)), b(256);
    // the tube mask exists in separate instantiations of the stage kernels
    const bool tube = tm.fx != nullptr;
#define LSMP_ADVK(M, S, TB)                                                                                     \
    hipLaunchKernelGGL((k_adv_twist<T, M, S, TB>), g4, b, 0, st, phi, vel, R, L1, L2, scr, nlines, n, lw, ostride, \
                       istride, estride, dd, dt, lines, tm)
#define LSMP_ADV(M, S)                                                                                          \
    do {                                                                                                        \
        if constexpr (M == 2) {                                                                                 \
            if (tube) LSMP_ADVK(M, S, true);                                                                    \
            else LSMP_ADVK(M, S, false);                                                                        \
        } else {                                                                                                \
            LSMP_ADVK(M, S, false);                                                                             \
        }                                                                                                       \
    } while (0)
    const bool part8 = n <= 32 * 8 + 1;   // eight rows per lane suffice
    const dim3 gp((unsigned)(((size_t)nlines * 64 + 255) / 256));
#define LSMP_PARTK(M, S, KR, TB)                                                                              \
    hipLaunchKernelGGL((k_adv_part<T, M, S, KR, TB>), gp, b, 0, st, phi, vel, R, L1, L2, nlines, n, lw, ostride, \
                       istride, estride, dd, dt, lines, tm)
#define LSMP_PART(M, S)                                                                                       \
    do {                                                                                                      \
        if constexpr (M == 2) {                                                                               \
            if (part8 && tube) LSMP_PARTK(M, S, 8, true);                                                     \
            else if (part8) LSMP_PARTK(M, S, 8, false);                                                       \
            else if (tube) LSMP_PARTK(M, S, PKX, true);                                                       \
            else LSMP_PARTK(M, S, PKX, false);                                                                \
        } else {                                                                                              \
            if (part8) LSMP_PARTK(M, S, 8, false);                                                            \
            else LSMP_PARTK(M, S, PKX, false);                                                                \
        }                                                                                                     \
    } while (0)
    if (mode == 0) {
        if (part) LSMP_PART(0, 0);
        else hipLaunchKernelGGL((k_adv_pair_x<T>), g, b, 0, st, phi, vel, R, scr, nlines, n, dd);
    } else if (part) {
        if (mode == 1) LSMP_PART(1, 1);
        else if (stage == 1) LSMP_PART(2, 1);
        else if (stage == 2) LSMP_PART(2, 2);
        else LSMP_PART(2, 3);
    } else if (mode == 1) LSMP_ADV(1, 1);
    else if (stage == 1) LSMP_ADV(2, 1);
    else if (stage == 2) LSMP_ADV(2, 2);
    else LSMP_ADV(2, 3);
#undef LSMP_ADV
#undef LSMP_ADVK
#undef LSMP_PART
#undef LSMP_PARTK
}

// Narrow band: flag the reinitialisation workgroups (R3X x R3Y x RKC_NB cells)
// holding a cell with |phi| < wband.  One workgroup per flag: thread (x, y)
// reads its column's RKC_NB cells, and one thread stores the flag, 1 or 0.
static_assert(R3X * R3Y == 256, "k_tube_tiles maps one thread to each column of a workgroup's cells");
template <typename T>
__global__ __launch_bounds__(256) void k_tube_tiles(const T *__restrict__ phi, int nx, int ny, int nz, double wband,
                                                    unsigned char *__restrict__ tiles)
{
    const int i = blockIdx.x * R3X + (threadIdx.x % R3X), j = blockIdx.y * R3Y + (threadIdx.x / R3X);
    const int k0 = blockIdx.z * RKC_NB;
    bool hit = false;
    if (i < nx && j < ny) {
        T v[RKC_NB];
#pragma unroll
        for (int kk = 0; kk < RKC_NB; ++kk)
            v[kk] = k0 + kk < nz ? phi[((size_t)(k0 + kk) * ny + j) * nx + i] : T(wband);
#pragma unroll
        for (int kk = 0; kk < RKC_NB; ++kk) hit = hit || fabs((double)v[kk]) < wband;
    }
    hit = __syncthreads_or(hit);
    if (threadIdx.x == 0)
        tiles[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = hit ? 1 : 0;
}

template <typename T>
void launch_tube_tiles(hipStream_t st, const T *phi, int nx, int ny, int nz, double wband, unsigned char *tiles)
{
    const dim3 g(nblk(nx, R3X), nblk(ny, R3Y), nblk(nz, RKC_NB));
    hipLaunchKernelGGL((k_tube_tiles<T>), g, dim3(256), 0, st, phi, nx, ny, nz, wband, tiles);
}

template <typename T>
void launch_redis_stage2d(hipStream_t st, int stage, int btn, const T *phi, T *phi_out, const T *phi0, const T *S0,
                          const RedisArgs &ra, unsigned long long *ctl)
{
    const dim3 g(nblk(ra.nx, RTX), nblk(ra.ny, RTY), 1);
#define LSMP_R2(S, B) hipLaunchKernelGGL((k_redis_stage2d<T, S, B>), g, dim3(RNT), 0, st, phi, phi_out, phi0, S0, ra, ctl)
    if (btn) { if (stage == 1) LSMP_R2(1, 1); else if (stage == 2) LSMP_R2(2, 1); else LSMP_R2(3, 1); }
    else { if (stage == 1) LSMP_R2(1, 0); else if (stage == 2) LSMP_R2(2, 0); else LSMP_R2(3, 0); }
#undef LSMP_R2
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

// lines == nullptr: every line of direction dir; else the nlist listed lines
// (a final pass over a list then sums the listed lines only).
template <typename T>
void launch_ccd_pass(hipStream_t st, int dir, int dim, bool final_pass, const T *phi, const T *coef, T *scr, T *out,
                     const T *dxv, const T *dyv, T *gr, double *partial, int nx, int ny, int nz, const CcdArgs &ca,
                     int *nblocks_out, const int *lines, int nlist)
{
    const int n = dir == 0 ? nx : dir == 1 ? ny : nz;
    const int nall = dir == 0 ? ny * nz : dir == 1 ? nx * nz : nx * ny;
    const int nlines = lines ? nlist : nall;
    const int lw = dir == 0 ? nall : nx;
    const long ostride = dir == 0 ? 0 : dir == 1 ? (long)nx * ny : (long)nx;
    const long istride = dir == 0 ? (long)nx : 1;
    const long estride = dir == 0 ? 1 : dir == 1 ? (long)nx : (long)nx * ny;
    const dim3 g(nblk(nlines, 256)), b(256);
    const size_t tab = (size_t)13 * (n + 6) * sizeof(T);
    const bool lds = tab <= 64 * 1024;
    const size_t shm = lds ? tab : 0;
#define LSMP_CCD(D, F)                                                                                         \
    do {                                                                                                       \
        if (lds)                                                                                               \
            hipLaunchKernelGGL((k_ccd_seg<T, D, F, true>), g, b, shm, st, phi, coef, scr, out, dxv, dyv, gr,     \
                               partial, nlines, n, lw, ostride, istride, estride, ca, lines);                  \
        else                                                                                                   \
            hipLaunchKernelGGL((k_ccd_seg<T, D, F, false>), g, b, 0, st, phi, coef, scr, out, dxv, dyv, gr,      \
                               partial, nlines, n, lw, ostride, istride, estride, ca, lines);                  \
    } while (0)
    // few lines (the 2D grids, or the lines crossing the band): 32 lanes per
    // line instead of one
    const int part_max = !lines ? PART_MAX_LINES_STRIDED : dir == 0 ? PART_MAX_LISTED_X : PART_MAX_LISTED_STRIDED;
    // lines of up to 540 cells: 32 lanes per line; up to 1084 cells: 64
    const bool part32 = n + 4 <= 32 * KCC;
    const bool part = nlines <= part_max && n + 4 <= 64 * KCC;
    const bool part9 = n + 4 <= 32 * 9;
    // two lines per wave (one with 64 lanes per line), four waves per block
    const dim3 gp(nblk(nlines, part32 ? 8 : 4));
#define LSMP_CCDQ(D, F, KC, LPL)                                                                               \
    do {                                                                                                       \
        if (lds)                                                                                               \
            hipLaunchKernelGGL((k_ccd_part<T, D, F, true, KC, LPL>), gp, b, shm, st, phi, coef, out, dxv, dyv,   \
                               gr, partial, nlines, n, lw, ostride, istride, estride, ca, lines);              \
        else                                                                                                   \
            hipLaunchKernelGGL((k_ccd_part<T, D, F, false, KC, LPL>), gp, b, 0, st, phi, coef, out, dxv, dyv,    \
                               gr, partial, nlines, n, lw, ostride, istride, estride, ca, lines);              \
    } while (0)
#define LSMP_CCDP(D, F)                                                                                        \
    do {                                                                                                       \
        if (part9) LSMP_CCDQ(D, F, 9, 32);                                                                     \
        else if (part32) LSMP_CCDQ(D, F, KCC, 32);                                                             \
        else LSMP_CCDQ(D, F, KCC, 64);                                                                         \
    } while (0)
    int nbl = 0;
    if (nlines > 0) {
        if (part) {
            if (!final_pass) LSMP_CCDP(3, false);
            else if (dim == 3) LSMP_CCDP(3, true);
            else LSMP_CCDP(2, true);
        } else if (!final_pass) LSMP_CCD(3, false);
        else if (dim == 3) LSMP_CCD(3, true);
        else LSMP_CCD(2, true);
        nbl = (int)(part ? gp.x : g.x);
    }
#undef LSMP_CCD
#undef LSMP_CCDP
#undef LSMP_CCDQ
    if (nblocks_out) *nblocks_out = final_pass ? nbl : 0;
}

template <typename T>
void launch_line_hsums(hipStream_t st, int dir, const T *phi, const unsigned char *lflag, int nx, int ny, int nz,
                       const CcdArgs &ca, double *partial, int *nblocks_out)
{
    const int n = dir == 0 ? nx : dir == 1 ? ny : nz;
    const int nall = dir == 0 ? ny * nz : dir == 1 ? nx * nz : nx * ny;
    const int lw = dir == 0 ? nall : nx;
    const long ostride = dir == 0 ? 0 : dir == 1 ? (long)nx * ny : (long)nx;
    const long istride = dir == 0 ? (long)nx : 1;
    const long estride = dir == 0 ? 1 : dir == 1 ? (long)nx : (long)nx * ny;
    // up to LINE_SEGMENTS segments of at least 32 cells per line
    const int nseg = std::max(1, std::min(LINE_SEGMENTS, n / 32));
    const int seg = (n + nseg - 1) / nseg;
    const dim3 g(nblk(nall, 256), (unsigned)nseg);
    hipLaunchKernelGGL((k_line_hsums<T>), g, dim3(256), 0, st, phi, lflag, nall, n, lw, ostride, istride, estride,
                       seg, ca, partial);
    *nblocks_out = (int)(g.x * g.y);
}

template <typename T>
void launch_mass_update_lines(hipStream_t st, int dir, T *phi, const T *gr, T ls_wid, const double *scal,
                              const int *lines, int nlist, int nx, int ny, int nz)
{
    if (nlist <= 0) return;
    const int n = dir == 0 ? nx : dir == 1 ? ny : nz;
    const int lw = dir == 0 ? ny * nz : nx;
    const long ostride = dir == 0 ? 0 : dir == 1 ? (long)nx * ny : (long)nx;
    const long istride = dir == 0 ? (long)nx : 1;
    const long estride = dir == 0 ? 1 : dir == 1 ? (long)nx : (long)nx * ny;
    const size_t total = (size_t)nlist * n;
    hipLaunchKernelGGL((k_mass_update_lines<T>), dim3(nblk(total, 256)), dim3(256), 0, st, phi, gr, ls_wid, scal,
                       lines, nlist, n, lw, ostride, istride, estride);
}

template <typename T>
void launch_uccd_pass(hipStream_t st, int dir, int mode, int stage, const T *phi, T *phi_out, const T *vel, T *R,
                      T *L1, T *L2, const T *coef, T *ck, int nx, int ny, int nz, T dx, T dt)
{
    const int n = dir == 0 ? nx : dir == 1 ? ny : nz;
    const int nlines = dir == 0 ? ny * nz : dir == 1 ? nx * nz : nx * ny;
    const int lw = dir == 0 ? nlines : nx;
    const long ostride = dir == 0 ? 0 : dir == 1 ? (long)nx * ny : (long)nx;
    const long istride = dir == 0 ? (long)nx : 1;
    const long estride = dir == 0 ? 1 : dir == 1 ? (long)nx : (long)nx * ny;
    const dim3 g(nblk(nlines, 256)), b(256);
    const size_t tab = (size_t)26 * (n + 6) * sizeof(T);
    const bool lds = tab <= 64 * 1024;
#define LSMP_UCCD(M, S)                                                                                        \
    do {                                                                                                       \
        if (lds)                                                                                               \
            hipLaunchKernelGGL((k_uccd_seg<T, M, S, true>), g, b, tab, st, phi, phi_out, vel, R, L1, L2, coef,   \
                               ck, nlines, n, lw, ostride, istride, estride, dx, dt);                          \
        else                                                                                                   \
            hipLaunchKernelGGL((k_uccd_seg<T, M, S, false>), g, b, 0, st, phi, phi_out, vel, R, L1, L2, coef,    \
                               ck, nlines, n, lw, ostride, istride, estride, dx, dt);                          \
    } while (0)
    if (mode == 0) LSMP_UCCD(0, 1);
    else if (mode == 1) LSMP_UCCD(1, 1);
    else if (stage == 1) LSMP_UCCD(2, 1);
    else if (stage == 2) LSMP_UCCD(2, 2);
    else LSMP_UCCD(2, 3);
#undef LSMP_UCCD
}

void launch_finalize3(hipStream_t st, const double *partial, int nblk_, double imass, double *scal)
{
    hipLaunchKernelGGL(k_finalize3, dim3(1), dim3(1024), 0, st, partial, nblk_, imass, scal);
}

template <typename T>
void launch_band_lines(hipStream_t st, int dim, T *phi, int nx, int ny, int nz, double wband,
                       unsigned char *flags, unsigned *count, int *listx, int *listy, int *listz, bool clamp)
{
    const int nfx = ny * nz, nfy = nx * nz, nfz = dim == 3 ? nx * ny : 0;
    const size_t nbytes = band_flag_bytes(nfx, nfy, nfz);
    hipLaunchKernelGGL(k_zero_bytes, dim3(nblk(nbytes, 256)), dim3(256), 0, st, flags, nbytes);
    const size_t n = (size_t)nx * ny * nz;
    unsigned char *fy = flags + band_flag_offset(1, nfx, nfy), *fz = flags + band_flag_offset(2, nfx, nfy);
    if (clamp)
        hipLaunchKernelGGL((k_band_line_flags<T, true>), dim3(nblk(n, 256)), dim3(256), 0, st, phi, nx, ny, nz,
                           wband, flags, fy, fz);
    else
        hipLaunchKernelGGL((k_band_line_flags<T, false>), dim3(nblk(n, 256)), dim3(256), 0, st, phi, nx, ny, nz,
                           wband, flags, fy, fz);
    hipLaunchKernelGGL(k_compact_lines, dim3(dim), dim3(1024), 0, st, flags, nfx, nfy, nfz, listx, listy, listz,
                       count);
}

void launch_zero_bytes(hipStream_t st, unsigned char *p, size_t n)
{
    hipLaunchKernelGGL(k_zero_bytes, dim3(nblk(n, 256)), dim3(256), 0, st, p, n);
}

void launch_compact_lines(hipStream_t st, int dim, const unsigned char *flags, int nfx, int nfy, int nfz,
                          int *listx, int *listy, int *listz, unsigned *count)
{
    hipLaunchKernelGGL(k_compact_lines, dim3(dim), dim3(1024), 0, st, flags, nfx, nfy, nfz, listx, listy, listz,
                       count);
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
void launch_velocity(hipStream_t st, T *u, T *v, T *w, const VelArgs &va, double *tab)
{
    const size_t n = (size_t)va.nx * va.ny * va.nz;
    if (va.field != 0)
        hipLaunchKernelGGL(k_vel_tables, dim3(nblk((size_t)va.nx + va.ny + va.nz, 256)), dim3(256), 0, st, va, tab);
    // rows of up to 256 cells, several rows per workgroup when they are short
    const int bx = (int)std::min<size_t>(256, (va.nx + 63) / 64 * 64), by = 256 / bx;
    const dim3 g((unsigned)((va.nx + bx - 1) / bx), (unsigned)std::min((va.ny + by - 1) / by, 65535),
                 (unsigned)std::min(va.nz, 65535));
    if (n) hipLaunchKernelGGL((k_velocity<T>), g, dim3(bx, by), 0, st, u, v, w, va, tab);
}

// Merge of the two markers of the MPLS two-marker branch
// (interface_solver.f90:85-111): where both have the same sign the merged phi
// is the smaller inside both and the larger outside both, elsewhere the
// larger; flag gets 1 when some cell is inside both.  The tests run in FP64
// as the reference's do (a product of two FP32 values is exact in FP64).
template <typename T>
__global__ void k_combine_markers(const T *__restrict__ m1, const T *__restrict__ m2, T *__restrict__ phi,
                                  size_t n, unsigned *__restrict__ flag)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;   // whole waves return together past the end (q grows with the lane)
    const double a = (double)m1[q], b = (double)m2[q];
    bool both = false;
    double out;
    if (a * b > 0.0) {
        both = a + b > 0.0;
        out = both ? fmin(a, b) : fmax(a, b);
    } else {
        out = fmax(a, b);
    }
    phi[q] = (T)out;
    if (__any(both) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

template <typename T>
void launch_combine_markers(hipStream_t st, const T *m1, const T *m2, T *phi, size_t n, unsigned *flag)
{
    hipLaunchKernelGGL((k_combine_markers<T>), dim3(nblk(n, 256)), dim3(256), 0, st, m1, m2, phi, n, flag);
}

template <typename TO, typename TI>
void launch_convert(hipStream_t st, TO *out, const TI *in, size_t n)
{
    hipLaunchKernelGGL((k_convert<TO, TI>), dim3(nblk(n, 256)), dim3(256), 0, st, out, in, n);
}

void launch_redis_ctl(hipStream_t st, unsigned long long *ctl, int op)
{
    static_assert(CTL_SLOTS == 64, "one wave, a stripe per lane");
    hipLaunchKernelGGL(k_redis_ctl, dim3(1), dim3(CTL_SLOTS), 0, st, ctl, op);
}

void launch_zero_u64(hipStream_t st, unsigned long long *p, int n)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, n);
}

#define LSMP_INST(T)                                                                                           \
    template void launch_adv_pass<T>(hipStream_t, int, int, int, T *, const T *, T *, T *, T *, T *, int, int, int, \
                                     T, T, const int *, int, const TubeMask &);                                \
    template void launch_combine_markers<T>(hipStream_t, const T *, const T *, T *, size_t, unsigned *);         \
    template void launch_redis_stage2d<T>(hipStream_t, int, int, const T *, T *, const T *, const T *,             \
                                          const RedisArgs &, unsigned long long *);                            \
    template void launch_tube_tiles<T>(hipStream_t, const T *, int, int, int, double, unsigned char *);          \
    template void launch_sign_init<T>(hipStream_t, const T *, T *, size_t, T, int);                            \
    template void launch_grad_max<T>(hipStream_t, int, const T *, const T *, const RedisArgs &,                \
                                     unsigned long long *);                                                    \
    template void launch_scale_by_max<T>(hipStream_t, T *, size_t, const unsigned long long *);                \
    template void launch_ccd_pass<T>(hipStream_t, int, int, bool, const T *, const T *, T *, T *, const T *,     \
                                     const T *, T *, double *, int, int, int, const CcdArgs &, int *,          \
                                     const int *, int);                                                        \
    template void launch_line_hsums<T>(hipStream_t, int, const T *, const unsigned char *, int, int, int,          \
                                       const CcdArgs &, double *, int *);                                      \
    template void launch_mass_update_lines<T>(hipStream_t, int, T *, const T *, T, const double *, const int *,   \
                                              int, int, int, int);                                             \
    template void launch_band_lines<T>(hipStream_t, int, T *, int, int, int, double, unsigned char *,            \
                                       unsigned *, int *, int *, int *, bool);                                 \
    template void launch_mass_update<T>(hipStream_t, T *, const T *, size_t, T, const double *);               \
    template void launch_mass_vol<T>(hipStream_t, const T *, size_t, const CcdArgs &, double *, int);           \
    template void launch_l1<T>(hipStream_t, const T *, const T *, size_t, double, double *, int);              \
    template void launch_velocity<T>(hipStream_t, T *, T *, T *, const VelArgs &, double *);                  \
    template void launch_uccd_pass<T>(hipStream_t, int, int, int, const T *, T *, const T *, T *, T *, T *,     \
                                      const T *, T *, int, int, int, T, T);

LSMP_INST(double)
LSMP_INST(float)
template void launch_convert<double, float>(hipStream_t, double *, const float *, size_t);
template void launch_convert<float, double>(hipStream_t, float *, const double *, size_t);

}  // namespace lsmp
