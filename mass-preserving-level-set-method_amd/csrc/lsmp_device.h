// lsmp_device.h -- per-point numerics of the level-set transport step, shared
// by the gfx950 kernels.  Each function restates one reference formula; the
// association order of every sum and product is the reference's (Fortran
// evaluates left to right; x**2.0d0 is a square).
#pragma once
#include <hip/hip_runtime.h>

#define LSMP_DEV __device__ __forceinline__

namespace lsmp {

template <typename T> LSMP_DEV T sq(T x) { return x * x; }
template <typename T> LSMP_DEV T dabs(T x) { return fabs(x); }
LSMP_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Workgroups are handed to the 8 XCDs of the device in turn, each XCD with its
// own L2.  This renumbers workgroup b so that XCD x takes a contiguous range
// (the first nb % 8 XCDs one workgroup more): neighbouring strided lines, which
// share cache lines, then run on the same XCD at about the same time.
LSMP_DEV int xcd_renumber(int b, int nb)
{
    const int x = b % 8, i = b / 8, q = nb / 8, r = nb % 8;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}
LSMP_DEV int xcd_block() { return xcd_renumber((int)blockIdx.x, (int)gridDim.x); }

// Pointwise flux split g = 1/2 (u + |u|) phi, f = 1/2 (u - |u|) phi
// (level_set_rk3_solver.f90:118-125).
template <typename T> LSMP_DEV void flux_split(T phi, T u, T &g, T &f)
{
    g = T(0.5) * (u + fabs(u)) * phi;
    f = T(0.5) * (u - fabs(u)) * phi;
}

// WENO-JS5 reconstruction of the positive flux at i+1/2 from g(i-2..i+2)
// (level_set_rk3_solver.f90:296-317).
template <typename T> LSMP_DEV T wenojs_g(T g0, T g1, T g2, T g3, T g4)
{
    const T EPS = T(1.0e-13);
    T b1 = T(13) * sq(g0 - T(2) * g1 + g2) + T(3) * sq(g0 - T(4) * g1 + T(3) * g2);
    T b2 = T(13) * sq(g1 - T(2) * g2 + g3) + T(3) * sq(g1 - g3);
    T b3 = T(13) * sq(g2 - T(2) * g3 + g4) + T(3) * sq(T(3) * g2 - T(4) * g3 + g4);
    T a1 = T(1) / sq(EPS + b1);
    T a2 = T(6) / sq(EPS + b2);
    T a3 = T(3) / sq(EPS + b3);
    T w1 = a1 / (a1 + a2 + a3);
    T w2 = a2 / (a1 + a2 + a3);
    T w3 = a3 / (a1 + a2 + a3);
    return w1 / T(3) * g0 - (T(7) * w1 + w2) / T(6) * g1 +
           (T(11) * w1 + T(5) * w2 + T(2) * w3) / T(6) * g2 +
           (T(2) * w2 + T(5) * w3) / T(6) * g3 - w3 / T(6) * g4;
}

// WENO-JS5 reconstruction of the negative flux at i+1/2 from f(i-1..i+3)
// (level_set_rk3_solver.f90:319-342).
template <typename T> LSMP_DEV T wenojs_f(T f0, T f1, T f2, T f3, T f4)
{
    const T EPS = T(1.0e-13);
    T b3 = T(13) * sq(f0 - T(2) * f1 + f2) + T(3) * sq(f0 - T(4) * f1 + T(3) * f2);
    T b2 = T(13) * sq(f1 - T(2) * f2 + f3) + T(3) * sq(f1 - f3);
    T b1 = T(13) * sq(f2 - T(2) * f3 + f4) + T(3) * sq(T(3) * f2 - T(4) * f3 + f4);
    T a1 = T(1) / sq(EPS + b1);
    T a2 = T(6) / sq(EPS + b2);
    T a3 = T(3) / sq(EPS + b3);
    T w1 = a1 / (a1 + a2 + a3);
    T w2 = a2 / (a1 + a2 + a3);
    T w3 = a3 / (a1 + a2 + a3);
    return w3 * (-f0 + T(5) * f1 + T(2) * f2) / T(6) +
           w2 * (T(2) * f1 + T(5) * f2 - f3) / T(6) +
           w1 * (T(11) * f2 - T(7) * f3 + T(2) * f4) / T(6);
}

// One CRWENO5 row of the positive-flux system from g(i-2..i+2)
// (level_set_rk3_solver.f90:362-381).
template <typename T>
LSMP_DEV void crweno_g_row(T g0, T g1, T g2, T g3, T g4, T &A, T &B, T &C, T &S)
{
    const T EPS = T(1.0e-13);
    const T c1 = T(0.2089141306), c2 = T(0.4999999998), c3 = T(0.2910858692);
    T b1 = T(13) * sq(g0 - T(2) * g1 + g2) + T(3) * sq(g0 - T(4) * g1 + T(3) * g2);
    T b2 = T(13) * sq(g1 - T(2) * g2 + g3) + T(3) * sq(g1 - g3);
    T b3 = T(13) * sq(g2 - T(2) * g3 + g4) + T(3) * sq(T(3) * g2 - T(4) * g3 + g4);
    T a1 = c1 * (T(1) + fabs(b3 - b1) / (EPS + b1));
    T a2 = c2 * (T(1) + fabs(b3 - b1) / (EPS + b2));
    T a3 = c3 * (T(1) + fabs(b3 - b1) / (EPS + b3));
    T w1 = a1 / (a1 + a2 + a3);
    T w2 = a2 / (a1 + a2 + a3);
    T w3 = a3 / (a1 + a2 + a3);
    A = (T(2) * w1 + w2) / T(3);
    B = (w1 + T(2) * (w2 + w3)) / T(3);
    C = w3 / T(3);
    S = w1 / T(6) * g1 + (T(5) * (w1 + w2) + w3) / T(6) * g2 + (w2 + T(5) * w3) / T(6) * g3;
}

// One CRWENO5 row of the negative-flux system from f(i-1..i+3)
// (level_set_rk3_solver.f90:388-407).
template <typename T>
LSMP_DEV void crweno_f_row(T f0, T f1, T f2, T f3, T f4, T &A, T &B, T &C, T &S)
{
    const T EPS = T(1.0e-13);
    const T c1 = T(0.2089141306), c2 = T(0.4999999998), c3 = T(0.2910858692);
    T b3 = T(13) * sq(f0 - T(2) * f1 + f2) + T(3) * sq(f0 - T(4) * f1 + T(3) * f2);
    T b2 = T(13) * sq(f1 - T(2) * f2 + f3) + T(3) * sq(f1 - f3);
    T b1 = T(13) * sq(f2 - T(2) * f3 + f4) + T(3) * sq(T(3) * f2 - T(4) * f3 + f4);
    T a1 = c1 * (T(1) + fabs(b3 - b1) / (EPS + b1));
    T a2 = c2 * (T(1) + fabs(b3 - b1) / (EPS + b2));
    T a3 = c3 * (T(1) + fabs(b3 - b1) / (EPS + b3));
    T w1 = a1 / (a1 + a2 + a3);
    T w2 = a2 / (a1 + a2 + a3);
    T w3 = a3 / (a1 + a2 + a3);
    A = (w3) / T(3);
    B = (w1 + T(2) * (w2 + w3)) / T(3);
    C = (w2 + T(2) * w1) / T(3);
    S = (T(5) * w3 + w2) / T(6) * f1 + (w3 + T(5) * (w2 + w1)) / T(6) * f2 + w1 / T(6) * f3;
}

// Reciprocal of a normal nonzero number (either sign): the hardware estimate
// refined by Newton steps (two in FP64, one in FP32), within an ulp or two of
// the correctly rounded 1/x.
LSMP_DEV double rcp_pos(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
LSMP_DEV float rcp_pos(float x)
{
    float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// The same with one Newton step in FP64: the hardware estimate is good to
// about 5e-8, one step to about 2e-15 of 1/x -- enough for the WENO-HJ
// weights, which multiply a correction term that is itself small.
LSMP_DEV double rcp_w(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}
LSMP_DEV float rcp_w(float x) { return rcp_pos(x); }

// sqrt of x >= 0 from the hardware reciprocal square root (good to about 6e-8)
// and one Newton step: good to about 5e-15 of sqrt(x); exactly 0 at x = 0.
LSMP_DEV double sqrt_w(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double s = x * y;
    const double r = fma(fma(-s, s, x), 0.5 * y, s);
    return x > 0.0 ? r : 0.0;
}
LSMP_DEV float sqrt_w(float x) { return __builtin_amdgcn_sqrtf(x); }

// The same two rows with the weight divisions folded into one reciprocal:
// with e_k = EPS + b_k and tau = |b3 - b1| the reference's
// a_k = c_k (1 + tau / e_k) normalise to
// w_k = c_k (e_k + tau) prod_{m != k} e_m / sum_j c_j (e_j + tau) prod_{m != j} e_m,
// and the divisions by 3 and 6 become products.  Equal to crweno_*_row() up to
// rounding of the weights.  The fold needs the FP64 exponent range: e_k lies in
// [1e-13, ~2e3], so a product of three can reach 1e-39, below FP32's normal
// range; the FP32 form keeps one reciprocal per weight.
LSMP_DEV void crweno_weights_fast(double b1, double b2, double b3, double &w1, double &w2, double &w3)
{
    const double EPS = 1.0e-13;
    const double c1 = 0.2089141306, c2 = 0.4999999998, c3 = 0.2910858692;
    const double tau = fabs(b3 - b1);
    const double e1 = EPS + b1, e2 = EPS + b2, e3 = EPS + b3;
    const double a1 = c1 * (e1 + tau) * (e2 * e3);
    const double a2 = c2 * (e2 + tau) * (e1 * e3);
    const double a3 = c3 * (e3 + tau) * (e1 * e2);
    const double r = rcp_pos(a1 + a2 + a3);
    w1 = a1 * r;
    w2 = a2 * r;
    w3 = a3 * r;
}
LSMP_DEV void crweno_weights_fast(float b1, float b2, float b3, float &w1, float &w2, float &w3)
{
    const float EPS = 1.0e-13f;
    const float c1 = 0.2089141306f, c2 = 0.4999999998f, c3 = 0.2910858692f;
    const float tau = fabsf(b3 - b1);
    const float a1 = c1 * (1.0f + tau * rcp_pos(EPS + b1));
    const float a2 = c2 * (1.0f + tau * rcp_pos(EPS + b2));
    const float a3 = c3 * (1.0f + tau * rcp_pos(EPS + b3));
    const float r = rcp_pos(a1 + a2 + a3);
    w1 = a1 * r;
    w2 = a2 * r;
    w3 = a3 * r;
}

template <typename T>
LSMP_DEV void crweno_g_row_fast(T g0, T g1, T g2, T g3, T g4, T &A, T &B, T &C, T &S)
{
    T b1 = T(13) * sq(g0 - T(2) * g1 + g2) + T(3) * sq(g0 - T(4) * g1 + T(3) * g2);
    T b2 = T(13) * sq(g1 - T(2) * g2 + g3) + T(3) * sq(g1 - g3);
    T b3 = T(13) * sq(g2 - T(2) * g3 + g4) + T(3) * sq(T(3) * g2 - T(4) * g3 + g4);
    T w1, w2, w3;
    crweno_weights_fast(b1, b2, b3, w1, w2, w3);
    const T third = T(1.0 / 3.0), sixth = T(1.0 / 6.0);
    A = (T(2) * w1 + w2) * third;
    B = (w1 + T(2) * (w2 + w3)) * third;
    C = w3 * third;
    S = (w1 * g1 + (T(5) * (w1 + w2) + w3) * g2 + (w2 + T(5) * w3) * g3) * sixth;
}

template <typename T>
LSMP_DEV void crweno_f_row_fast(T f0, T f1, T f2, T f3, T f4, T &A, T &B, T &C, T &S)
{
    T b3 = T(13) * sq(f0 - T(2) * f1 + f2) + T(3) * sq(f0 - T(4) * f1 + T(3) * f2);
    T b2 = T(13) * sq(f1 - T(2) * f2 + f3) + T(3) * sq(f1 - f3);
    T b1 = T(13) * sq(f2 - T(2) * f3 + f4) + T(3) * sq(T(3) * f2 - T(4) * f3 + f4);
    T w1, w2, w3;
    crweno_weights_fast(b1, b2, b3, w1, w2, w3);
    const T third = T(1.0 / 3.0), sixth = T(1.0 / 6.0);
    A = w3 * third;
    B = (w1 + T(2) * (w2 + w3)) * third;
    C = (w2 + T(2) * w1) * third;
    S = ((T(5) * w3 + w2) * f1 + (w3 + T(5) * (w2 + w1)) * f2 + w1 * f3) * sixth;
}

// WENO-HJ correction term (level_set_rk3_redis.f90:456-471).
template <typename T> LSMP_DEV T phyn(T a, T b, T c, T d)
{
    const T eps = T(1.0e-6);
    T is0 = T(13) * sq(a - b) + T(3) * sq(a - T(3) * b);
    T is1 = T(13) * sq(b - c) + T(3) * sq(b + c);
    T is2 = T(13) * sq(c - d) + T(3) * sq(T(3) * c - d);
    T alp0 = T(1) / sq(eps + is0);
    T alp1 = T(6) / sq(eps + is1);
    T alp2 = T(3) / sq(eps + is2);
    T w0 = alp0 / (alp0 + alp1 + alp2);
    T w2 = alp2 / (alp0 + alp1 + alp2);
    return w0 / T(3) * (a - T(2) * b + c) + (w2 - T(0.5)) / T(6) * (b - T(2) * c + d);
}

// The same WENO-HJ term with its five divisions folded into one reciprocal:
// with q_k = (eps + IS_k)^2 the reference's alp0 = 1/q0, alp1 = 6/q1,
// alp2 = 3/q2 normalise to w0 = q1 q2 / D and w2 = 3 q0 q1 / D with
// D = q1 q2 + 6 q0 q2 + 3 q0 q1.  Equal to phyn() up to rounding of the weights
// (q_k lies in [1e-12, 3e6], so D neither under- nor overflows in FP32).
// sbc = (b - c)^2 and scd = (c - d)^2 come from the caller: the two terms of
// one direction share them (squares of opposite differences are equal).
template <typename T> LSMP_DEV T phyn_fast(T a, T b, T c, T d, T sbc, T scd)
{
    const T eps = T(1.0e-6);
    T is0 = T(13) * sq(a - b) + T(3) * sq(a - T(3) * b);
    T is1 = T(13) * sbc + T(3) * sq(b + c);
    T is2 = T(13) * scd + T(3) * sq(T(3) * c - d);
    T q0 = sq(eps + is0), q1 = sq(eps + is1), q2 = sq(eps + is2);
    T q12 = q1 * q2, q01 = q0 * q1;
    T r = rcp_w(q12 + T(6) * (q0 * q2) + T(3) * q01);
    // w0/3 = q12 r/3 and (w2 - 1/2)/6 = q01 r/2 - 1/12
    T t0 = q12 * (r * T(1.0 / 3.0));
    T t2 = fma(q01, r * T(0.5), T(-1.0 / 12.0));
    return fma(t0, a - T(2) * b + c, t2 * (b - T(2) * c + d));
}

// Face m+1/2 of a line with second differences sa = s(m-1), sb = s(m),
// sc = s(m+1), sd = s(m+2): plus = 12 phyn(sd, sc, sb, sa) (the term of D+
// at m) and minus = 12 phyn(sa, sb, sc, sd) (the term of D- at m+1), each as
// phyn_fast evaluates it up to rounding; the factor 12 joins the one of the
// central part, so that both scale by 1 / (12 dx).  IS0 of one is IS2 of the other,
// IS1 is common, and so are the products of their squares and the two
// third differences the terms weigh.  The indicators are the expanded
// quadratic forms 13 (a - b)^2 + 3 (a - 3b)^2 = 4 (4a^2 - 11ab + 10b^2) and
// 13 (a - b)^2 + 3 (a + b)^2 = 4 (4a^2 - 5ab + 4b^2) over shared squares,
// each formed as 4 IS + eps = 40b^2 + (a (16a - 44b) + eps) (and
// 16a^2 + (b (16b - 20a) + eps)).  With the weights' denominators Dp and
// Dm, plus = (4 pJB e1 + 6 pAJ e2) / Dp - e2 and minus the same with the
// roles of the two products and third differences exchanged.
// FP64 takes both reciprocals from one (the product of the two denominators
// stays within [1e-48, 1e28]).  three: 3.0.
template <typename T> LSMP_DEV void weno_hj_face(T sa, T sb, T sc, T sd, T three, T &plus, T &minus)
{
    const T eps = T(1.0e-6);
    const T bb = sb * sb, cc = sc * sc;
    const T fA = fma(T(40), cc, fma(sd, fma(T(16), sd, T(-44) * sc), eps));   // IS0 of plus, IS2 of minus
    const T fJ = fma(T(16), bb, fma(sc, fma(T(16), sc, T(-20) * sb), eps));   // IS1 of both
    const T fB = fma(T(40), bb, fma(sa, fma(T(16), sa, T(-44) * sb), eps));   // IS2 of plus, IS0 of minus
    const T qA = sq(fA), qJ = sq(fJ), qB = sq(fB);
    const T pJB = qJ * qB, pAJ = qA * qJ, pAB = qA * qB;
    // plus: q0 = qA, q1 = qJ, q2 = qB; minus: q0 = qB, q1 = qJ, q2 = qA
    const T dp = fma(three, fma(T(2), pAB, pAJ), pJB), dm = fma(three, fma(T(2), pAB, pJB), pAJ);
    T rp, rm;
    if constexpr (sizeof(T) == 8) {
        const T r = rcp_w(dp * dm);
        rp = dm * r;
        rm = dp * r;
    } else {
        rp = rcp_w(dp);
        rm = rcp_w(dm);
    }
    const T e1 = sd - T(2) * sc + sb, e2 = sc - T(2) * sb + sa;
    const T u1 = pJB * e1, u2 = pAJ * e2;
    plus = fma(rp, fma(T(4), u1, T(6) * u2), -e2);
    minus = fma(rm, fma(T(4), u2, T(6) * u1), -e1);
}

// The central part of the WENO-HJ derivatives times 12 dx,
// -(m1 - m2) + 7 (c - m1) + 7 (p1 - c) - (p2 - p1), which is
// (m2 - p2) + 8 (p1 - m1): c drops out.
template <typename T> LSMP_DEV T central12(T m2, T m1, T p1, T p2)
{
    return fma(T(8), p1 - m1, m2 - p2);
}

// D+ and D- from the five second differences s_m = phi(m+1) - 2 phi(m) + phi(m-1),
// m = i-2..i+2, shared by both one-sided terms.
template <typename T>
LSMP_DEV void weno_hj_pm_fast(T m3, T m2, T m1, T c, T p1, T p2, T p3, T inv12d, T invd, T &dp, T &dm)
{
    const T a = inv12d * (-(m1 - m2) + T(7) * (c - m1) + T(7) * (p1 - c) - (p2 - p1));
    const T sm2 = m3 - T(2) * m2 + m1, sm1 = m2 - T(2) * m1 + c, s0 = m1 - T(2) * c + p1;
    const T sp1 = c - T(2) * p1 + p2, sp2 = p1 - T(2) * p2 + p3;
    // (sp1 - s0)^2 is is1 of D+ and is2 of D-; (s0 - sm1)^2 is2 of D+ and is1 of D-
    const T qp = sq(sp1 - s0), qm = sq(s0 - sm1);
    dp = a + invd * phyn_fast(sp2, sp1, s0, sm1, qp, qm);
    dm = a - invd * phyn_fast(sm2, sm1, s0, sp1, qm, qp);
}

// ls_funs with the divisions by ls_wid and pi replaced by products with
// precomputed reciprocals.
template <typename T> LSMP_DEV void ls_funs_fast(T phi, T inv_w, T &H, T &D)
{
    const T eps = T(1.0e-12);
    const T pi = T(3.141592653589793238462643383279502884);
    const T inv_pi = T(0.318309886183790671537767526745028724);
    T x = phi * inv_w;
    if (x > T(1) - eps) {
        H = T(1); D = T(0);
    } else if (x < T(-1) + eps) {
        H = T(0); D = T(0);
    } else {
        H = T(0.5) * ((T(1) + x) + sin(pi * x) * inv_pi);
        D = T(0.5) * (T(1) + cos(pi * x)) * inv_w;
    }
    H = fmax(fmin(H, T(1)), T(0));
}

// Fourth-order central slope times 12 dx (level_set_rk3_redis.f90:290-293).
template <typename T> LSMP_DEV T central4(T m2, T m1, T c, T p1, T p2)
{
    return -(m1 - m2) + T(7) * (c - m1) + T(7) * (p1 - c) - (p2 - p1);
}

// D+ and D- along one direction from phi(i-3..i+3) (level_set_rk3_redis.f90:305-315).
template <typename T>
LSMP_DEV void weno_hj_pm(T m3, T m2, T m1, T c, T p1, T p2, T p3, T inv12d, T invd, T &dp, T &dm)
{
    T a = inv12d * central4(m2, m1, c, p1, p2);
    dp = a + invd * phyn(p3 - T(2) * p2 + p1, p2 - T(2) * p1 + c, p1 - T(2) * c + m1, c - T(2) * m1 + m2);
    dm = a - invd * phyn(m3 - T(2) * m2 + m1, m2 - T(2) * m1 + c, m1 - T(2) * c + p1, c - T(2) * p1 + p2);
}

// Godunov |grad phi| selected by the frozen sign S0 (level_set_rk3_redis.f90:341-360).
template <typename T, int DIM>
LSMP_DEV T godunov(T xs1, T xs2, T ys1, T ys2, T zs1, T zs2, T s0)
{
    T upm = -fmin(xs1, T(0)), upp = fmax(xs1, T(0));
    T umm = -fmin(xs2, T(0)), ump = fmax(xs2, T(0));
    T vpm = -fmin(ys1, T(0)), vpp = fmax(ys1, T(0));
    T vmm = -fmin(ys2, T(0)), vmp = fmax(ys2, T(0));
    if (DIM == 3) {
        T wpm = -fmin(zs1, T(0)), wpp = fmax(zs1, T(0));
        T wmm = -fmin(zs2, T(0)), wmp = fmax(zs2, T(0));
        if (s0 >= T(0)) return sqrt(sq(fmax(upm, ump)) + sq(fmax(vpm, vmp)) + sq(fmax(wpm, wmp)));
        return sqrt(sq(fmax(upp, umm)) + sq(fmax(vpp, vmm)) + sq(fmax(wpp, wmm)));
    } else {
        if (s0 >= T(0)) return sqrt(sq(fmax(upm, ump)) + sq(fmax(vpm, vmp)));
        return sqrt(sq(fmax(upp, umm)) + sq(fmax(vpp, vmm)));
    }
}

// Smoothed Heaviside and delta (ns_solver/others.f90:18-31).
template <typename T> LSMP_DEV void ls_funs(T phi, T ls_wid, T &H, T &D)
{
    const T eps = T(1.0e-12);
    const T pi = T(3.141592653589793238462643383279502884);
    T x = phi / ls_wid;
    if (x > T(1) - eps) {
        H = T(1); D = T(0);
    } else if (x < T(-1) + eps) {
        H = T(0); D = T(0);
    } else {
        H = T(0.5) * ((T(1) + x) + sin(pi * x) / pi);
        D = T(0.5) * (T(1) + cos(pi * x)) / ls_wid;
    }
    H = fmax(fmin(H, T(1)), T(0));
}

// Heaviside used by the shape metric, without the clip (output.f90:233-248).
template <typename T> LSMP_DEV T heavyside(T x)
{
    const T eps = T(1.0e-12);
    const T pi = T(3.141592653589793238462643383279502884);
    if (x > T(1) - eps) return T(1);
    if (x < T(-1) + eps) return T(0);
    return T(0.5) * ((T(1) + x) + sin(pi * x) / pi);
}

// Wave- and block-wide reductions (64-lane wavefronts).
LSMP_DEV double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
LSMP_DEV double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide sum of NV doubles; result valid in thread 0.  blockDim.x must be
// a multiple of 64 and at most 1024.
template <int NV> LSMP_DEV void block_sum(double (&v)[NV])
{
    __shared__ double red[16][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0)
        for (int q = 0; q < NV; ++q) red[wid][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < NV; ++q) {
            double s = red[0][q];
            for (int w = 1; w < nw; ++w) s += red[w][q];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Neumaier-compensated FP64 summation: s carries the rounded running sum and
// c the rounding errors of its additions (each recovered exactly), so s + c is
// the sum to about one rounding whatever the number of terms.
LSMP_DEV void csum_add(double &s, double &c, double x)
{
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

// Block-wide compensated sum of NV (s, c) pairs (blockDim.x a multiple of 64,
// at most 1024): pairs combine by adding their sums with csum_add and their
// compensations plainly; thread 0 gets s[q] = s + c of the block.
template <int NV> LSMP_DEV void block_sum_comp(double (&s)[NV], double (&c)[NV])
{
    __shared__ double red_s[16][NV], red_c[16][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q)
        for (int o = 32; o > 0; o >>= 1) {
            const double s2 = __shfl_xor(s[q], o, 64), c2 = __shfl_xor(c[q], o, 64);
            c[q] += c2;
            csum_add(s[q], c[q], s2);
        }
    if (lane == 0)
        for (int q = 0; q < NV; ++q) {
            red_s[wid][q] = s[q];
            red_c[wid][q] = c[q];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < NV; ++q) {
            double ss = red_s[0][q], cc = red_c[0][q];
            for (int w = 1; w < nw; ++w) {
                cc += red_c[w][q];
                csum_add(ss, cc, red_s[w][q]);
            }
            s[q] = ss + cc;
        }
    }
    __syncthreads();
}

// Compensated sum of the per-block partial triples partial[3 b + q], b < nblk,
// over the whole block (slot b to thread b, b + blockDim.x, ...); the totals
// land in acc of thread 0.  Fixed order: the result does not depend on timing.
LSMP_DEV void sum_partials3(const double *__restrict__ partial, int nblk, double (&acc)[3])
{
    // four slots' loads in flight, then their additions in slot order
    constexpr int AHEAD = 4;
    double cmp[3] = {0.0, 0.0, 0.0};
    acc[0] = acc[1] = acc[2] = 0.0;
    const int bd = blockDim.x;
    for (int b0 = threadIdx.x; b0 < nblk; b0 += AHEAD * bd) {
        double v[AHEAD][3];
#pragma unroll
        for (int u = 0; u < AHEAD; ++u) {
            const int b = b0 + u * bd;
#pragma unroll
            for (int q = 0; q < 3; ++q) v[u][q] = b < nblk ? partial[3 * b + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < AHEAD; ++u)
            if (b0 + u * bd < nblk)
#pragma unroll
                for (int q = 0; q < 3; ++q) csum_add(acc[q], cmp[q], v[u][q]);
    }
    block_sum_comp<3>(acc, cmp);
}

// max of non-negative doubles through their bit patterns (ordered as integers)
LSMP_DEV void atomic_max_nonneg(unsigned long long *p, double v)
{
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

LSMP_DEV double block_max(double v)
{
    __shared__ double red[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    v = wave_max(v);
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < nw; ++w) v = fmax(v, red[w]);
    __syncthreads();
    return v;
}

}  // namespace lsmp
