// lsmp_learn.hip -- learned reinitialisation: a 128-64-64-1 SiLU MLP in BF16
// on MFMA (FP32 accumulation) that maps a narrow-band 5^3 phi stencil to a
// signed-distance correction, trained on the GPU from (pre, post) pairs of
// the plain reinitialisation, then used in place of most pseudo-steps.
//
// This part has no counterpart in the reference; it replaces most of the
// pseudo-time iterations of level_set_rk3_redis (level_set_rk3_redis.f90:25-74)
// and hands the result to the plain method for a few refining sweeps.
//
// MFMA: __builtin_amdgcn_mfma_f32_16x16x32_bf16, per lane: A row = lane%16,
// A/B k = 8*(lane/16) .. +7; B column = lane%16; D row = 4*(lane/16)+i,
// D column = lane%16 (i = 0..3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lsmp_ctx.h"
#include "lsmp_device.h"

namespace lsmp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int NF = 128;      // 125 stencil values, the centre value, 2 zeros
constexpr int NH = 64;
constexpr int OFF_W1 = 0;                    // W1[k][n], k < NF, n < NH
constexpr int OFF_B1 = OFF_W1 + NF * NH;
constexpr int OFF_W2 = OFF_B1 + NH;          // W2[k][n], k, n < NH
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;          // W3[k], k < NH
constexpr int OFF_B3 = OFF_W3 + NH;
constexpr int NPARAM = OFF_B3 + 1;           // 12481
constexpr int TILE = 64;                     // samples per workgroup (4 waves x 16)
// LDS row pitches (in bf16) of the MFMA operand tiles: 16 bytes past a
// multiple of the bank width, so that the 16 rows a lane group reads at once
// fall on different banks
constexpr int KP1 = NF + 8, KP2 = NH + 8, KPT = TILE + 8;

LSMP_DEV f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
// the logistic function from the hardware 2^x and reciprocal (each within
// about an ulp): SiLU to a few ulp, far finer than the BF16 products it feeds
LSMP_DEV float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z)); }
LSMP_DEV float silu(float z) { return z * sigm(z); }
LSMP_DEV float dsilu(float z)
{
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}
LSMP_DEV float group16_sum(float v)   // sum over the 16 lanes sharing lane/16
{
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}
LSMP_DEV float across_groups_sum(float v)   // sum over the 4 lanes sharing lane%16
{
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

struct Geom {
    int nx, ny, nz;
    float inv_h;      // 1 / dx
    float inv_12h;    // 1 / (12 dx)
};

template <typename T> LSMP_DEV T load_clamped(const T *__restrict__ phi, const Geom &g, int a, int b, int c)
{
    a = clampi(a, 0, g.nx - 1);
    b = clampi(b, 0, g.ny - 1);
    c = clampi(c, 0, g.nz - 1);
    return phi[((size_t)c * g.ny + b) * g.nx + a];
}

// |grad phi| at the cell from fourth-order central differences (clamped indices).
template <typename T> LSMP_DEV float grad_norm4(const T *__restrict__ phi, const Geom &g, int i, int j, int k)
{
    auto d4 = [&](int di, int dj, int dk) {
        const T m2 = load_clamped(phi, g, i - 2 * di, j - 2 * dj, k - 2 * dk);
        const T m1 = load_clamped(phi, g, i - di, j - dj, k - dk);
        const T p1 = load_clamped(phi, g, i + di, j + dj, k + dk);
        const T p2 = load_clamped(phi, g, i + 2 * di, j + 2 * dj, k + 2 * dk);
        return (float)((m2 - p2) + T(8) * (p1 - m1)) * g.inv_12h;
    };
    const float gx = d4(1, 0, 0), gy = d4(0, 1, 0), gz = d4(0, 0, 1);
    return sqrtf(gx * gx + gy * gy + gz * gz);
}

// Feature q of the cell at (i, j, k):
//   q < 125:  (phi(nb) - phi_c) / h over the 5^3 stencil (x offset fastest,
//             clamped indices: the zero-gradient ghost rule);
//   q = 125:  phi_c / h;
//   q = 126:  |grad phi| - 1 (fourth-order central differences);
//   q = 127:  (phi_c / h) (1 / max(|grad phi|, 0.1) - 1), the change that
//             the first-order distance estimate phi / |grad phi| would make.
// gm: |grad phi| at the cell (grad_norm4), for features 126 and 127.
template <typename T>
LSMP_DEV float feature(const T *__restrict__ phi, const Geom &g, int i, int j, int k, T pc, int q, float gm)
{
    if (q < 125) {
        const int di = q % 5 - 2, dj = (q / 5) % 5 - 2, dk = q / 25 - 2;
        return (float)(load_clamped(phi, g, i + di, j + dj, k + dk) - pc) * g.inv_h;
    }
    if (q == 125) return (float)pc * g.inv_h;
    if (q == 126) return gm - 1.0f;
    return (float)pc * g.inv_h * (1.0f / fmaxf(gm, 0.1f) - 1.0f);
}
template <typename T>
LSMP_DEV float feature(const T *__restrict__ phi, const Geom &g, int i, int j, int k, T pc, int q)
{
    return feature(phi, g, i, j, k, pc, q, q >= 126 ? grad_norm4(phi, g, i, j, k) : 1.0f);
}

// ---------------------------------------------------------------------------
// band list, pair collection, residual
// ---------------------------------------------------------------------------
// Ordered narrow-band compaction: cells with |phi| < thr in ascending index
// order.  Each workgroup of 256 threads owns BAND_CHUNK consecutive cells as
// BAND_ROWS rows of 256 (thread t reads cell t of every row, so every load is
// one coalesced 1 KB (FP32) or 2 KB (FP64) row); pass 1 counts, a
// one-workgroup scan makes the workgroups' offsets, pass 2 places each cell by
// the ballots of its row: the cells of earlier rows, of earlier waves in its
// row, and of lower lanes in its wave come first.
constexpr int BAND_ROWS = 16;
constexpr int BAND_CHUNK = 256 * BAND_ROWS;
static_assert(4 * BAND_ROWS == 64, "k_band_write scans the 4 x BAND_ROWS wave counts with one wave");

template <typename T>
LSMP_DEV bool band_flag(const T *__restrict__ phi, size_t n, T thr, size_t q)
{
    return q < n && fabs(phi[q]) < thr;
}

// With S0 given, the pass also stores the sign field S0 = 2 H(phi) - 1 of
// every cell (k_sign_init's values), saving that kernel's read of phi.
template <typename T>
__global__ __launch_bounds__(256) void k_band_count(const T *__restrict__ phi, size_t n, T thr,
                                                    unsigned *__restrict__ cnt, T *__restrict__ S0, T ls_wid)
{
    __shared__ unsigned wsum[4];
    const size_t q0 = (size_t)blockIdx.x * BAND_CHUNK + threadIdx.x;
    bool f[BAND_ROWS];
    if (S0) {
#pragma unroll
        for (int r = 0; r < BAND_ROWS; ++r) {
            const size_t q = q0 + (size_t)r * 256;
            const T p = q < n ? phi[q] : T(0);
            f[r] = q < n && fabs(p) < thr;
            if (q < n) {
                T H, D;
                ls_funs(p, ls_wid, H, D);
                S0[q] = T(2) * H - T(1);
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < BAND_ROWS; ++r) f[r] = band_flag(phi, n, thr, q0 + (size_t)r * 256);
    }
    unsigned c = 0;
#pragma unroll
    for (int r = 0; r < BAND_ROWS; ++r) c += (unsigned)__popcll(__ballot(f[r]));
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Exclusive scan of cnt[0..nblk) into off[], total into off[nblk]; one
// workgroup of 1024 threads, each summing a contiguous run of counts, an
// exclusive scan of the runs' sums, then each run written from its offset.
__global__ __launch_bounds__(1024) void k_band_scan(const unsigned *__restrict__ cnt, int nblk,
                                                    unsigned *__restrict__ off)
{
    __shared__ unsigned wsum[16];
    const int per = (nblk + 1023) / 1024;
    const int b0 = min((int)threadIdx.x * per, nblk), b1 = min(b0 + per, nblk);
    unsigned run = 0;
    for (int b = b0; b < b1; ++b) run += cnt[b];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    unsigned inc = run;
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
        const unsigned y = __shfl_up(inc, k, 64);
        if (lane >= k) inc += y;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    unsigned base = 0;
    for (int w = 0; w < wid; ++w) base += wsum[w];
    unsigned o = base + inc - run;
    for (int b = b0; b < b1; ++b) {
        off[b] = o;
        o += cnt[b];
    }
    if (threadIdx.x == 1023) off[nblk] = base + inc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_band_write(const T *__restrict__ phi, size_t n, T thr,
                                                    const unsigned *__restrict__ off, int *__restrict__ list)
{
    // pre[4 r + w]: cells of row r's wave w, then their exclusive prefix in
    // (row, wave) order
    __shared__ unsigned pre[4 * BAND_ROWS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t q0 = (size_t)blockIdx.x * BAND_CHUNK + threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    bool f[BAND_ROWS];
    unsigned rank[BAND_ROWS];
#pragma unroll
    for (int r = 0; r < BAND_ROWS; ++r) f[r] = band_flag(phi, n, thr, q0 + (size_t)r * 256);
#pragma unroll
    for (int r = 0; r < BAND_ROWS; ++r) {
        const unsigned long long m = __ballot(f[r]);
        rank[r] = (unsigned)__popcll(m & below);
        if (lane == 0) pre[4 * r + w] = (unsigned)__popcll(m);
    }
    __syncthreads();
    if (w == 0) {
        const unsigned v = pre[lane];   // 4 * BAND_ROWS == 64 entries
        unsigned x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        pre[lane] = x - v;
    }
    __syncthreads();
    const unsigned base = off[blockIdx.x];
#pragma unroll
    for (int r = 0; r < BAND_ROWS; ++r)
        if (f[r]) list[base + pre[4 * r + w] + rank[r]] = (int)(q0 + (size_t)r * 256);
}

// Pair c takes band entry first + c * stride.
template <typename T>
__global__ void k_collect(const T *__restrict__ pre, const T *__restrict__ post, const int *__restrict__ list,
                          int nlist, int first, int stride, Geom g, float *__restrict__ X, float *__restrict__ Y,
                          unsigned base, unsigned cap)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nlist) return;
    const size_t s = (size_t)base + c;
    if (s >= cap) return;
    const int idx = list[first + (size_t)c * stride];
    const int i = idx % g.nx, j = (idx / g.nx) % g.ny, k = idx / (g.nx * g.ny);
    const T pc = pre[idx];
    for (int q = 0; q < NF; ++q) X[s * NF + q] = feature(pre, g, i, j, k, pc, q);
    Y[s] = (float)(post[idx] - pc) * g.inv_h;
}

// band-averaged | |grad phi| - 1 | with the Godunov gradient signed by S0.
// With ctl given, the field is the result of `launched` queued pseudo-steps
// whose stop flag may have skipped some: phi_alt when launched - ctl[2] (the
// steps run) is odd, phi_in otherwise.
template <typename T, int DIM>
__global__ void k_residual(const T *__restrict__ phi_in, const T *__restrict__ S0, RedisArgs ra, double thr,
                           double *__restrict__ partial, const T *__restrict__ phi_alt,
                           const unsigned long long *__restrict__ ctl, int launched)
{
    const T *__restrict__ phi = ctl && ((launched - (int)ctl[2]) & 1) ? phi_alt : phi_in;
    const int nx = ra.nx, ny = ra.ny, nz = ra.nz;
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < (size_t)nx * ny * nz;
         q += (size_t)gridDim.x * blockDim.x) {
        if (!(fabs((double)phi[q]) < thr)) continue;
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
        auto P = [&](int a, int b, int c) {
            return phi[((size_t)clampi(c, 0, nz - 1) * ny + clampi(b, 0, ny - 1)) * nx + clampi(a, 0, nx - 1)];
        };
        T xs1, xs2, ys1, ys2, zs1 = T(0), zs2 = T(0);
        weno_hj_pm_fast(P(i - 3, j, k), P(i - 2, j, k), P(i - 1, j, k), P(i, j, k), P(i + 1, j, k), P(i + 2, j, k),
                        P(i + 3, j, k), (T)ra.inv12dx, (T)ra.invdx, xs1, xs2);
        weno_hj_pm_fast(P(i, j - 3, k), P(i, j - 2, k), P(i, j - 1, k), P(i, j, k), P(i, j + 1, k), P(i, j + 2, k),
                        P(i, j + 3, k), (T)ra.inv12dy, (T)ra.invdy, ys1, ys2);
        if (DIM == 3)
            weno_hj_pm_fast(P(i, j, k - 3), P(i, j, k - 2), P(i, j, k - 1), P(i, j, k), P(i, j, k + 1),
                            P(i, j, k + 2), P(i, j, k + 3), (T)ra.inv12dz, (T)ra.invdz, zs1, zs2);
        const double g = (double)godunov<T, DIM>(xs1, xs2, ys1, ys2, zs1, zs2, S0[q]);
        acc[0] += fabs(g - 1.0);
        acc[1] += 1.0;
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = acc[0];
        partial[3 * blockIdx.x + 1] = acc[1];
        partial[3 * blockIdx.x + 2] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// network
// ---------------------------------------------------------------------------
// Split-BF16 products: an FP32 value v is carried as hi = bf16(v) and
// lo = bf16(v - hi) (v - hi is exact in FP32), and a product of two such
// operands is hi*hi + lo*hi + hi*lo, three BF16 MFMAs accumulated in FP32
// (the dropped lo*lo term is below 2^-16 relative).
LSMP_DEV void split_bf16(float v, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}

struct NetLds {
    __attribute__((aligned(16))) __bf16 w1h[NH][KP1];   // W1 transposed, [n][k], hi part
    __attribute__((aligned(16))) __bf16 w1l[NH][KP1];   // lo part
    __attribute__((aligned(16))) __bf16 w2h[NH][KP2];   // W2 transposed, [n2][n1]
    __attribute__((aligned(16))) __bf16 w2l[NH][KP2];
    float b1[NH], b2[NH], w3[NH];
    float b3;
};

// w2: optional [n1][n2] copy of the hi part of W2 for the backward pass.
LSMP_DEV void stage_weights(NetLds &L, const float *__restrict__ P, __bf16 (*w2)[KP2])
{
    for (int idx = threadIdx.x; idx < NF * NH; idx += blockDim.x) {
        const int k = idx / NH, n = idx % NH;
        split_bf16(P[OFF_W1 + idx], L.w1h[n][k], L.w1l[n][k]);
    }
    for (int idx = threadIdx.x; idx < NH * NH; idx += blockDim.x) {
        const int k = idx / NH, n = idx % NH;
        split_bf16(P[OFF_W2 + idx], L.w2h[n][k], L.w2l[n][k]);
        if (w2) w2[k][n] = L.w2h[n][k];
    }
    for (int idx = threadIdx.x; idx < NH; idx += blockDim.x) {
        L.b1[idx] = P[OFF_B1 + idx];
        L.b2[idx] = P[OFF_B2 + idx];
        L.w3[idx] = P[OFF_W3 + idx];
    }
    if (threadIdx.x == 0) L.b3 = P[OFF_B3];
}

// Forward pass of one wave's 16 samples.  xh/xl: the input rows in A layout
// (hi and lo parts).  Returns z1, z2 (pre-activations, D layout) and the
// outputs o[i] of rows 4*(lane/16)+i; sh/sl are this wave's 16x64 LDS tiles
// for the H1 relayout (hi and lo parts; sh is kept for the backward pass).
LSMP_DEV void net_forward(const NetLds &L, const bf16x8 (&xh)[4], const bf16x8 (&xl)[4], __bf16 (*sh)[KP2],
                          __bf16 (*sl)[KP2], f32x4 (&z1)[4], f32x4 (&z2)[4], float (&o)[4])
{
    const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const bf16x8 wh = *(const bf16x8 *)&L.w1h[16 * nt + lr][32 * ks + 8 * lg];
            const bf16x8 wl = *(const bf16x8 *)&L.w1l[16 * nt + lr][32 * ks + 8 * lg];
            acc = mma(xl[ks], wh, acc);
            acc = mma(xh[ks], wl, acc);
            acc = mma(xh[ks], wh, acc);
        }
        const float b = L.b1[16 * nt + lr];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            split_bf16(silu(acc[i]), sh[4 * lg + i][16 * nt + lr], sl[4 * lg + i][16 * nt + lr]);
        }
        z1[nt] = acc;
    }
    __syncthreads();
    bf16x8 ah[2], al[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        ah[ks] = *(const bf16x8 *)&sh[lr][32 * ks + 8 * lg];
        al[ks] = *(const bf16x8 *)&sl[lr][32 * ks + 8 * lg];
    }
    float p[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8 wh = *(const bf16x8 *)&L.w2h[16 * nt + lr][32 * ks + 8 * lg];
            const bf16x8 wl = *(const bf16x8 *)&L.w2l[16 * nt + lr][32 * ks + 8 * lg];
            acc = mma(al[ks], wh, acc);
            acc = mma(ah[ks], wl, acc);
            acc = mma(ah[ks], wh, acc);
        }
        const float b = L.b2[16 * nt + lr], w = L.w3[16 * nt + lr];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            p[i] += silu(acc[i]) * w;
        }
        z2[nt] = acc;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = group16_sum(p[i]) + L.b3;
}

// Rows of X for this wave in A layout, split into hi and lo parts.
LSMP_DEV void load_rows(const float *__restrict__ X, long srow, bool valid, bf16x8 (&xh)[4], bf16x8 (&xl)[4])
{
    const int l = threadIdx.x & 63, lg = l >> 4;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        float4 a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
        if (valid) {
            a = *(const float4 *)&X[srow * NF + 32 * ks + 8 * lg];
            b = *(const float4 *)&X[srow * NF + 32 * ks + 8 * lg + 4];
        }
        const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            __bf16 hi, lo;
            split_bf16(v[j], hi, lo);
            xh[ks][j] = hi;
            xl[ks][j] = lo;
        }
    }
}

LSMP_DEV unsigned hash32(unsigned x)
{
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// One training step's forward and backward pass over a minibatch: each
// workgroup handles 64 samples and writes its gradient of
// L = (1/B) sum w (o - y)^2, w = c^2 / (y^2 + c^2) (w = 1 for c = 0), to
// part[blockIdx.x][NPARAM] and its loss sum to lpart[blockIdx.x].  Samples
// are drawn with replacement by a hash of (seed, position), or taken in
// order when `ordered` is set.
__global__ __launch_bounds__(256) void k_net_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                  const float *__restrict__ Y, unsigned nsamp, int batch,
                                                  unsigned seed, int ordered, float wc, float *__restrict__ part,
                                                  double *__restrict__ lpart)
{
    __shared__ NetLds L;
    __shared__ __attribute__((aligned(16))) __bf16 w2[NH][KP2];        // [n1][n2]
    __shared__ __attribute__((aligned(16))) __bf16 sh1[4][16][KP2];     // per wave [row][n1], hi part
    __shared__ __attribute__((aligned(16))) __bf16 sl1[4][16][KP2];     // lo part
    __shared__ __attribute__((aligned(16))) __bf16 sdz2[4][16][KP2];    // per wave [row][n2]
    __shared__ __attribute__((aligned(16))) __bf16 h1t[NH][KPT];        // [n1][sample]
    __shared__ __attribute__((aligned(16))) __bf16 dz2t[NH][KPT];       // [n2][sample]
    __shared__ __attribute__((aligned(16))) __bf16 dz1t[NH][KPT];       // [n1][sample]
    __shared__ __attribute__((aligned(16))) __bf16 xt[NF][KPT];         // [feature][sample]
    __shared__ float red[4][3 * NH + 2];
    stage_weights(L, P, w2);
    __syncthreads();

    const int t = threadIdx.x, wv = t >> 6, l = t & 63, lr = l & 15, lg = l >> 4;
    const int pos = blockIdx.x * TILE + wv * 16 + lr;    // position in the minibatch of this lane's row
    const bool valid = pos < batch;
    unsigned s = 0;
    if (valid) s = ordered ? (unsigned)pos : hash32(seed * 0x9e3779b9u + (unsigned)pos) % nsamp;
    bf16x8 xh[4], xl[4];
    load_rows(X, (long)s, valid, xh, xl);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) xt[32 * ks + 8 * lg + j][wv * 16 + lr] = xh[ks][j];
    f32x4 z1[4], z2[4];
    float o[4];
    net_forward(L, xh, xl, sh1[wv], sl1[wv], z1, z2, o);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) h1t[16 * nt + lr][wv * 16 + 4 * lg + i] = sh1[wv][4 * lg + i][16 * nt + lr];

    // output layer: dO = 2 (o - y) / B for the rows 4*lg + i
    const float invB = 1.0f / (float)batch;
    float dO[4], lsum = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int rpos = blockIdx.x * TILE + wv * 16 + 4 * lg + i;
        float e = 0.f, w = 1.f;
        if (rpos < batch) {
            const unsigned rs = ordered ? (unsigned)rpos : hash32(seed * 0x9e3779b9u + (unsigned)rpos) % nsamp;
            const float y = Y[rs];
            e = o[i] - y;
            if (wc > 0.f) w = wc * wc / (y * y + wc * wc);
        }
        dO[i] = 2.0f * w * e * invB;
        lsum += w * e * e;
    }
    float gw3[4], gb2[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float w = L.w3[16 * nt + lr];
        float pw = 0.f, pb = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float h2 = silu(z2[nt][i]);
            pw += h2 * dO[i];
            const float d = dO[i] * w * dsilu(z2[nt][i]);
            pb += d;
            sdz2[wv][4 * lg + i][16 * nt + lr] = (__bf16)d;
            dz2t[16 * nt + lr][wv * 16 + 4 * lg + i] = (__bf16)d;
        }
        gw3[nt] = across_groups_sum(pw);
        gb2[nt] = across_groups_sum(pb);
    }
    float gb3 = across_groups_sum(dO[0] + dO[1] + dO[2] + dO[3]);
    __syncthreads();

    // dH1 = dZ2 W2^T, dZ1 = dH1 * silu'(Z1)
    bf16x8 ad[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) ad[ks] = *(const bf16x8 *)&sdz2[wv][lr][32 * ks + 8 * lg];
    float gb1[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) acc = mma(ad[ks], *(const bf16x8 *)&w2[16 * nt + lr][32 * ks + 8 * lg], acc);
        float pb = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float d = acc[i] * dsilu(z1[nt][i]);
            pb += d;
            dz1t[16 * nt + lr][wv * 16 + 4 * lg + i] = (__bf16)d;
        }
        gb1[nt] = across_groups_sum(pb);
    }
    // per-wave vector gradients into LDS, summed over waves below
    if (lg == 0) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            red[wv][16 * nt + lr] = gb1[nt];
            red[wv][NH + 16 * nt + lr] = gb2[nt];
            red[wv][2 * NH + 16 * nt + lr] = gw3[nt];
        }
    }
    const float lw = across_groups_sum(group16_sum(lr == 0 ? lsum : 0.f));   // all lanes take part
    if (l == 0) {
        red[wv][3 * NH] = gb3;
        red[wv][3 * NH + 1] = lw;
    }
    __syncthreads();

    float *out = part + (size_t)blockIdx.x * NPARAM;
    // dW2 = H1^T dZ2 (64 x 64, K = 64 samples): wave wv computes rows 16 wv ..
    {
        const int mt = wv;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = mma(*(const bf16x8 *)&h1t[16 * mt + lr][32 * ks + 8 * lg],
                          *(const bf16x8 *)&dz2t[16 * nt + lr][32 * ks + 8 * lg], acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) out[OFF_W2 + (16 * mt + 4 * lg + i) * NH + 16 * nt + lr] = acc[i];
        }
    }
    // dW1 = X^T dZ1 (128 x 64): wave wv computes feature rows 32 wv .. 32 wv + 31
#pragma unroll
    for (int mm = 0; mm < 2; ++mm) {
        const int mt = 2 * wv + mm;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = mma(*(const bf16x8 *)&xt[16 * mt + lr][32 * ks + 8 * lg],
                          *(const bf16x8 *)&dz1t[16 * nt + lr][32 * ks + 8 * lg], acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) out[OFF_W1 + (16 * mt + 4 * lg + i) * NH + 16 * nt + lr] = acc[i];
        }
    }
    for (int idx = t; idx < NH; idx += blockDim.x) {
        out[OFF_B1 + idx] = red[0][idx] + red[1][idx] + red[2][idx] + red[3][idx];
        out[OFF_B2 + idx] = red[0][NH + idx] + red[1][NH + idx] + red[2][NH + idx] + red[3][NH + idx];
        out[OFF_W3 + idx] = red[0][2 * NH + idx] + red[1][2 * NH + idx] + red[2][2 * NH + idx] + red[3][2 * NH + idx];
    }
    if (t == 0) {
        out[OFF_B3] = red[0][3 * NH] + red[1][3 * NH] + red[2][3 * NH] + red[3][3 * NH];
        lpart[blockIdx.x] = (double)red[0][3 * NH + 1] + red[1][3 * NH + 1] + red[2][3 * NH + 1] + red[3][3 * NH + 1];
    }
}

// Sum the per-workgroup gradients in fixed order; with adam != 0 apply one
// Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) to P.
__global__ void k_net_update(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                             const float *__restrict__ part, int nblk, float lr, float bc1, float bc2, int adam,
                             float *__restrict__ gsum, const double *__restrict__ lpart, double *__restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += lpart[b];
        *loss = s;
    }
    if (p >= NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += part[(size_t)b * NPARAM + p];
    if (gsum) gsum[p] = g;
    if (!adam) return;
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + 1.0e-8f);
}

// Inference on the band cells: val[e] = phi_c + h * net(features) * gate for
// band entry e (cell list[e]), gate = clamp((|phi_c| - w) / w, 0, 1) with
// w = ls_wid: the network leaves the interface region |phi| <= ls_wid to the
// plain sweeps and ramps in over [ls_wid, 2 ls_wid].  The band's length is
// read from device memory (*nlist_p).  Persistent workgroups: each stages the
// weights into LDS once and then takes tiles of TILE band cells blockIdx.x,
// + gridDim.x, ...; workgroups past the last tile return at once.
template <typename T>
__global__ __launch_bounds__(256) void k_net_infer(const T *__restrict__ phi, T *__restrict__ val,
                                                   const int *__restrict__ list,
                                                   const unsigned *__restrict__ nlist_p, Geom g,
                                                   const float *__restrict__ P, T h, T w)
{
    const int nlist = (int)*nlist_p;
    const int ntiles = (nlist + TILE - 1) / TILE;
    if ((int)blockIdx.x >= ntiles) return;   // the whole workgroup, before any barrier
    __shared__ NetLds L;
    __shared__ __attribute__((aligned(16))) __bf16 sh1[4][16][KP2];
    __shared__ __attribute__((aligned(16))) __bf16 sl1[4][16][KP2];
    __shared__ float sout[4][16];
    // stencil offset of feature q in cells, for cells 2 or more cells from
    // every face; features 125..127 take the stencil's last cell and do not use it
    __shared__ __attribute__((aligned(16))) int soff[NF];
    const long sy = g.nx, sz = (long)g.nx * g.ny;
    if (threadIdx.x < NF) {
        const int q = min((int)threadIdx.x, 124);
        soff[threadIdx.x] = (q % 5 - 2) + (int)sy * ((q / 5) % 5 - 2) + (int)sz * (q / 25 - 2);
    }
    stage_weights(L, P, nullptr);
    __syncthreads();
    const int t = threadIdx.x, wv = t >> 6, l = t & 63, lr = l & 15, lg = l >> 4;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int c = tile * TILE + wv * 16 + lr;
        const bool valid = c < nlist;
        // lanes past the list's end take the tile's first cell and store nothing
        const int idx = list[valid ? c : tile * TILE];
        const int i = idx % g.nx, j = (idx / g.nx) % g.ny, k = idx / (g.nx * g.ny);
        const T pc = phi[idx];
        // the 32 stencil values and the 12 of |grad phi|, all loads in flight
        // at once: from the offsets when no lane of the wave needs clamping,
        // else with clamped indices (the zero-gradient ghost rule)
        const bool inner = i >= 2 && i < g.nx - 2 && j >= 2 && j < g.ny - 2 && k >= 2 && k < g.nz - 2;
        T sv[32];
        float gm;
        if (__all(inner)) {
            const T *pcell = phi + idx;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) sv[8 * ks + jj] = pcell[soff[32 * ks + 8 * lg + jj]];
            auto d4 = [&](long st) {
                return (float)((pcell[-2 * st] - pcell[2 * st]) + T(8) * (pcell[st] - pcell[-st])) * g.inv_12h;
            };
            const float gx = d4(1), gy = d4(sy), gz = d4(sz);
            gm = sqrtf(gx * gx + gy * gy + gz * gz);
        } else {
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) {
                    const int q = min(32 * ks + 8 * lg + jj, 124);
                    sv[8 * ks + jj] = load_clamped(phi, g, i + q % 5 - 2, j + (q / 5) % 5 - 2, k + q / 25 - 2);
                }
            gm = grad_norm4(phi, g, i, j, k);
        }
        bf16x8 xh[4], xl[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const int q = 32 * ks + 8 * lg + jj;
                float f = (float)(sv[8 * ks + jj] - pc) * g.inv_h;   // feature() for q < 125
                if (q == 125) f = (float)pc * g.inv_h;
                else if (q == 126) f = gm - 1.0f;
                else if (q == 127) f = (float)pc * g.inv_h * (1.0f / fmaxf(gm, 0.1f) - 1.0f);
                __bf16 hi, lo;
                split_bf16(valid ? f : 0.0f, hi, lo);
                xh[ks][jj] = hi;
                xl[ks][jj] = lo;
            }
        f32x4 z1[4], z2[4];
        float o[4];
        net_forward(L, xh, xl, sh1[wv], sl1[wv], z1, z2, o);
        if (lr == 0)
            for (int q = 0; q < 4; ++q) sout[wv][4 * lg + q] = o[q];
        __syncthreads();
        if (valid && lg == 0) {
            const T gate = fmin(fmax((fabs(pc) - w) / w, T(0)), T(1));
            val[c] = pc + h * (T)sout[wv][lr] * gate;
        }
        __syncthreads();   // sout and the wave's H1 tiles are reused by the next tile
    }
}

// phi[list[e]] = val[e] for the *nlist_p band entries (the network's values,
// formed from the whole field before any is stored)
template <typename T>
__global__ __launch_bounds__(256) void k_band_scatter(T *__restrict__ phi, const T *__restrict__ val,
                                                      const int *__restrict__ list,
                                                      const unsigned *__restrict__ nlist_p)
{
    const size_t n = *nlist_p;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        phi[list[e]] = val[e];
}

template <typename TO, typename TI>
__global__ void k_copy_cast(TO *__restrict__ out, const TI *__restrict__ in, size_t n)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = (TO)in[q];
}

__global__ void k_init_params(float *__restrict__ P, unsigned seed)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    // uniform(-1, 1) scaled by sqrt(3 / fan_in): unit-variance pre-activations
    const float u = (float)(hash32(seed ^ (0x85ebca6bu * (unsigned)(p + 1))) >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
    float v = 0.0f;
    if (p < OFF_B1) v = u * sqrtf(3.0f / 126.0f);
    else if (p >= OFF_W2 && p < OFF_B2) v = u * sqrtf(3.0f / NH);
    else if (p >= OFF_W3 && p < OFF_B3) v = u * sqrtf(3.0f / NH) * 0.1f;
    P[p] = v;
}

}  // namespace lsmp

using namespace lsmp;

// ======================================================================
// host side
// ======================================================================
struct lsmp_net {
    hipStream_t st = nullptr;
    float *P = nullptr, *m = nullptr, *v = nullptr, *part = nullptr, *gsum = nullptr;
    double *lpart = nullptr, *loss = nullptr;
    float *X = nullptr;   // training inputs [capacity][NF]
    float *Y = nullptr;
    unsigned capacity = 0, nsamp = 0;
    int part_blocks = 0;
    int step = 0;
    double *hbuf = nullptr;
};

namespace {

thread_local char g_nerr[256];

#define NET_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            snprintf(g_nerr, sizeof g_nerr, "%s: %s", #call, hipGetErrorString(e_));               \
            fprintf(stderr, "lsmp: %s\n", g_nerr);                                                 \
            return LSMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

unsigned cdiv(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

int ensure_parts(lsmp_net *net, int nblk)
{
    if (nblk <= net->part_blocks) return LSMP_OK;
    if (net->part) (void)hipFree(net->part);
    if (net->lpart) (void)hipFree(net->lpart);
    net->part = nullptr;
    net->lpart = nullptr;
    NET_HIP(hipMalloc(&net->part, (size_t)nblk * NPARAM * sizeof(float)));
    NET_HIP(hipMalloc(&net->lpart, (size_t)nblk * sizeof(double)));
    net->part_blocks = nblk;
    return LSMP_OK;
}

int ensure_band_list(lsmp_ctx *c)
{
    if (c->band_list) return LSMP_OK;
    const size_t nblk = (c->ncell + BAND_CHUNK - 1) / BAND_CHUNK;
    NET_HIP(hipMalloc(&c->band_list, c->ncell * sizeof(int)));
    NET_HIP(hipMalloc(&c->band_count, (2 * nblk + 1) * sizeof(unsigned)));
    NET_HIP(hipMalloc(&c->band_val, c->ncell * (c->fp32 ? sizeof(float) : sizeof(double))));
    return LSMP_OK;
}

// Queues the ordered list of the cells with |phi| < thr; *count_dev is where
// its length lands in device memory.  S0, when given, receives 2 H(phi) - 1.
template <typename T>
int queue_band(lsmp_ctx *c, const T *phi, double thr, const unsigned **count_dev, T *S0 = nullptr)
{
    int rc = ensure_band_list(c);
    if (rc) return rc;
    const int nblk = (int)((c->ncell + BAND_CHUNK - 1) / BAND_CHUNK);
    unsigned *cnt = c->band_count, *off = c->band_count + nblk;
    hipLaunchKernelGGL((k_band_count<T>), dim3(nblk), dim3(256), 0, c->st, phi, c->ncell, (T)thr, cnt, S0,
                       (T)c->p.ls_wid);
    hipLaunchKernelGGL(k_band_scan, dim3(1), dim3(1024), 0, c->st, cnt, nblk, off);
    hipLaunchKernelGGL((k_band_write<T>), dim3(nblk), dim3(256), 0, c->st, phi, c->ncell, (T)thr, off,
                       c->band_list);
    NET_HIP(hipGetLastError());
    *count_dev = off + nblk;
    return LSMP_OK;
}

// The same list; returns its length.
template <typename T> int build_band(lsmp_ctx *c, const T *phi, double thr, unsigned *n_out)
{
    const unsigned *count_dev = nullptr;
    int rc = queue_band<T>(c, phi, thr, &count_dev);
    if (rc) return rc;
    NET_HIP(hipMemcpyAsync(c->hbuf, count_dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
    NET_HIP(hipStreamSynchronize(c->st));
    unsigned n = 0;
    std::memcpy(&n, c->hbuf, sizeof n);
    *n_out = n;
    return LSMP_OK;
}

Geom geom(const lsmp_ctx *c)
{
    Geom g;
    g.nx = c->nx; g.ny = c->ny; g.nz = c->nz;
    g.inv_h = (float)(1.0 / c->p.dx);
    g.inv_12h = (float)(1.0 / (12.0 * c->p.dx));
    return g;
}

template <typename T> int collect(lsmp_net *net, lsmp_ctx *c, double band_cells, int stride, int *n_added)
{
    const size_t n = c->ncell;
    T *pre = (T *)c->R;   // the advection RHS is free between steps
    NET_HIP(hipMemcpyAsync(pre, c->phi, n * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    unsigned nb = 0;
    int rc = build_band<T>(c, pre, band_cells * c->p.dx, &nb);
    if (rc) return rc;
    double err;
    int np;
    rc = lsmp_reinit_run(c, 1, -1.0, 0, true, &err, &np);
    if (rc) return rc;
    // every stride-th band cell, starting at an offset that varies between calls
    const unsigned first = nb ? net->nsamp % (unsigned)stride % nb : 0;
    const unsigned avail = nb > first ? (nb - first + stride - 1) / stride : 0;
    const unsigned room = net->capacity - net->nsamp;
    const unsigned take = avail < room ? avail : room;
    if (take) {
        hipLaunchKernelGGL((k_collect<T>), dim3(cdiv(take, 256)), dim3(256), 0, c->st, pre, (const T *)c->phi,
                           c->band_list, (int)take, (int)first, stride, geom(c), net->X, net->Y, net->nsamp,
                           net->capacity);
        NET_HIP(hipGetLastError());
    }
    NET_HIP(hipStreamSynchronize(c->st));
    net->nsamp += take;
    if (n_added) *n_added = (int)take;
    return LSMP_OK;
}

// pending > 0: the pseudo-steps of lsmp_reinit_queue are still open; the
// residual reads the buffer that holds their result, and its read-back also
// brings the control block, which closes them (err, np as lsmp_reinit_run).
template <typename T>
int residual(lsmp_ctx *c, double band_cells, double *res, int pending = 0, double *err = nullptr, int *np = nullptr)
{
    const RedisArgs ra = lsmp_redis_args(c);
    // as many workgroups as there are partial slots (10,240 at 256^3): the
    // grid-stride loop is load-latency bound and needs the waves
    size_t b = (c->ncell + 255) / 256;
    if ((int)b > c->npartial) b = c->npartial;
    const unsigned long long *ctl = pending > 0 ? c->rctl : nullptr;
    if (c->dim == 3)
        hipLaunchKernelGGL((k_residual<T, 3>), dim3((unsigned)b), dim3(256), 0, c->st, (const T *)c->phi,
                           (const T *)c->S0, ra, band_cells * c->p.dx, c->partial, (const T *)c->phiB, ctl,
                           pending);
    else
        hipLaunchKernelGGL((k_residual<T, 2>), dim3((unsigned)b), dim3(256), 0, c->st, (const T *)c->phi,
                           (const T *)c->S0, ra, band_cells * c->p.dx, c->partial, (const T *)c->phiB, ctl,
                           pending);
    NET_HIP(hipGetLastError());
    launch_finalize3(c->st, c->partial, (int)b, 0.0, c->scal);
    NET_HIP(hipGetLastError());
    NET_HIP(hipMemcpyAsync(c->hbuf, c->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    if (ctl)
        NET_HIP(hipMemcpyAsync(c->hbuf + 2, ctl, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
    NET_HIP(hipStreamSynchronize(c->st));
    *res = c->hbuf[1] > 0 ? c->hbuf[0] / c->hbuf[1] : 0.0;
    if (ctl) {
        unsigned long long hc[4];
        std::memcpy(hc, c->hbuf + 2, sizeof hc);
        const int rc = lsmp_reinit_finish(c, pending, hc, err, np);
        if (rc) return rc;
    }
    return LSMP_OK;
}

// phi <- network(phi) on |phi| < band_cells dx (phi unchanged elsewhere), in
// place: the network's values go to band_val, then onto phi.  No host round
// trip unless the band's length is asked for.  S0, when given, receives the
// sign field 2 H(phi) - 1 of the input (from the band list's counting pass).
template <typename T> int apply(lsmp_ctx *c, lsmp_net *net, double band_cells, int *n_band, T *S0 = nullptr)
{
    T *phi = (T *)c->phi;
    const unsigned *nb_dev = nullptr;
    int rc = queue_band<T>(c, phi, band_cells * c->p.dx, &nb_dev, S0);
    if (rc) return rc;
    T *val = (T *)c->band_val;   // allocated with the list on the first call
    // persistent workgroups, two per CU (the weights fill a third of the LDS)
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const unsigned grid = std::min<unsigned>(cdiv(c->ncell, TILE), 2u * (unsigned)std::max(ncu, 1));
    hipLaunchKernelGGL((k_net_infer<T>), dim3(grid), dim3(256), 0, c->st, phi, val, c->band_list, nb_dev, geom(c),
                       net->P, (T)c->p.dx, (T)c->p.ls_wid);
    hipLaunchKernelGGL((k_band_scatter<T>), dim3(std::min<unsigned>(cdiv(c->ncell, 256), 1024u)), dim3(256), 0,
                       c->st, phi, val, c->band_list, nb_dev);
    NET_HIP(hipGetLastError());
    if (n_band) {
        NET_HIP(hipMemcpyAsync(c->hbuf, nb_dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        NET_HIP(hipStreamSynchronize(c->st));
        unsigned nb = 0;
        std::memcpy(&nb, c->hbuf, sizeof nb);
        *n_band = (int)nb;
    }
    return LSMP_OK;
}

// Pseudo-steps the plain reinitialisation runs without an early exit: the
// loop of level_set_rk3_redis advances time by rdt and stops once time > timestop.
int plain_budget(const lsmp_ctx *c)
{
    const double timestop = 2.0 * c->p.ls_wid;
    double t = 0.0;
    int k = 0;
    do {
        ++k;
        t += c->p.rdt;
    } while (!(t > timestop) && k < 100000);
    return k;
}

template <typename T>
int reinit_learned(lsmp_ctx *c, lsmp_net *net, double band_cells, int sweeps, double tol, double res_band, int mode,
                   double *res_out, int *fallback, int *n_pseudo)
{
    const size_t n = c->ncell;
    T *phi = (T *)c->phi, *backup = (T *)c->R, *S0 = (T *)c->S0;
    // the sign of the field before the correction, as level_set_redis_init
    // sets it: stored by the network pass's band counting (which reads phi
    // before any change), or by its own kernel in the control without the
    // network (net == nullptr)
    if (!net) {
        launch_sign_init<T>(c->st, phi, S0, n, (T)c->p.ls_wid, 0);
        NET_HIP(hipGetLastError());
    }
    if (mode == LSMP_FALLBACK_RESTART)
        NET_HIP(hipMemcpyAsync(backup, phi, n * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    int rc = net ? apply<T>(c, net, band_cells, nullptr, S0) : LSMP_OK;
    if (rc) return rc;
    double err = 0.0;
    int np = 0;
    int queued = 0;
    rc = lsmp_reinit_queue(c, 1, sweeps, false, &queued, &err, &np);
    if (rc) return rc;
    double res = 0.0;
    // the residual's read-back also closes the queued sweeps
    rc = residual<T>(c, res_band, &res, queued, &err, &np);
    if (rc) return rc;
    *fallback = 0;
    if (!(res <= tol)) {
        *fallback = 1;
        if (mode == LSMP_FALLBACK_RESTART) {
            // the network's result is rejected: the plain method from the input field
            NET_HIP(hipMemcpyAsync(c->phi, backup, n * sizeof(T), hipMemcpyDeviceToDevice, c->st));
            int np2 = 0;
            rc = lsmp_reinit_run(c, 1, -1.0, 0, true, &err, &np2);
            if (rc) return rc;
            np += np2;
            rc = residual<T>(c, res_band, &res);
            if (rc) return rc;
        } else {
            // more plain pseudo-steps from the current field, `sweeps` at a time,
            // until the residual passes, the plain method's own stop (max change
            // <= 1e-6) is met, or its pseudo-time budget is spent
            // (CONTINUE_RELATIVE also stops when a group lowers the residual by
            // less than 10 %: what the plain method can reach is reached)
            int budget = plain_budget(c) - np;
            bool progress = true;
            while (budget > 0 && !(res <= tol) && err > 1.0e-6 && progress) {
                int np2 = 0, queued2 = 0;
                rc = lsmp_reinit_queue(c, 1, sweeps < budget ? sweeps : budget, false, &queued2, &err, &np2);
                if (rc) return rc;
                const double prev = res;
                rc = residual<T>(c, res_band, &res, queued2, &err, &np2);
                if (rc) return rc;
                np += np2;
                budget -= np2;
                if (mode == LSMP_FALLBACK_CONTINUE_RELATIVE) progress = res < 0.9 * prev;
            }
        }
    }
    *res_out = res;
    *n_pseudo = np;
    return LSMP_OK;
}

}  // namespace

extern "C" {

int lsmp_net_create(int capacity, unsigned seed, lsmp_net_handle *out)
{
    if (!out || capacity < 0) return LSMP_E_INVALID;
    *out = nullptr;
    lsmp_net *net = new lsmp_net();
    auto fail = [&](int code) {
        lsmp_net_destroy(net);
        return code;
    };
    if (hipStreamCreateWithFlags(&net->st, hipStreamNonBlocking) != hipSuccess) return fail(LSMP_E_HIP);
    if (hipMalloc(&net->P, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&net->m, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&net->v, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&net->gsum, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&net->loss, sizeof(double)) != hipSuccess)
        return fail(LSMP_E_HIP);
    net->capacity = (unsigned)capacity;
    if (capacity > 0) {
        if (hipMalloc(&net->X, (size_t)capacity * NF * sizeof(float)) != hipSuccess ||
            hipMalloc(&net->Y, (size_t)capacity * sizeof(float)) != hipSuccess)
            return fail(LSMP_E_HIP);
    }
    if (hipHostMalloc((void **)&net->hbuf, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(LSMP_E_HIP);
    hipLaunchKernelGGL(k_init_params, dim3(cdiv(NPARAM, 256)), dim3(256), 0, net->st, net->P, seed);
    if (hipGetLastError() != hipSuccess || hipMemsetAsync(net->m, 0, NPARAM * sizeof(float), net->st) != hipSuccess ||
        hipMemsetAsync(net->v, 0, NPARAM * sizeof(float), net->st) != hipSuccess ||
        hipStreamSynchronize(net->st) != hipSuccess)
        return fail(LSMP_E_HIP);
    *out = net;
    return LSMP_OK;
}

void lsmp_net_destroy(lsmp_net_handle net)
{
    if (!net) return;
    if (net->st) (void)hipStreamSynchronize(net->st);
    void *bufs[] = {net->P, net->m, net->v, net->part, net->gsum, net->lpart, net->loss, net->X, net->Y};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (net->hbuf) (void)hipHostFree(net->hbuf);
    if (net->st) (void)hipStreamDestroy(net->st);
    delete net;
}

int lsmp_net_param_count(void) { return NPARAM; }

int lsmp_net_set_params(lsmp_net_handle net, const float *params, int n)
{
    if (!net || !params || n != NPARAM) return LSMP_E_INVALID;
    NET_HIP(hipMemcpyAsync(net->P, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, net->st));
    NET_HIP(hipMemsetAsync(net->m, 0, NPARAM * sizeof(float), net->st));
    NET_HIP(hipMemsetAsync(net->v, 0, NPARAM * sizeof(float), net->st));
    NET_HIP(hipStreamSynchronize(net->st));
    net->step = 0;
    return LSMP_OK;
}

int lsmp_net_get_params(lsmp_net_handle net, float *params, int n)
{
    if (!net || !params || n != NPARAM) return LSMP_E_INVALID;
    NET_HIP(hipMemcpyAsync(params, net->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->st));
    NET_HIP(hipStreamSynchronize(net->st));
    return LSMP_OK;
}

int lsmp_net_dataset_size(lsmp_net_handle net, int *n)
{
    if (!net || !n) return LSMP_E_INVALID;
    *n = (int)net->nsamp;
    return LSMP_OK;
}

int lsmp_net_collect(lsmp_net_handle net, lsmp_handle h, double band_cells, int stride, int *n_added)
{
    if (!net || !h || !(band_cells > 0) || stride < 1 || net->capacity == 0) return LSMP_E_INVALID;
    return h->fp32 ? collect<float>(net, h, band_cells, stride, n_added)
                   : collect<double>(net, h, band_cells, stride, n_added);
}

int lsmp_net_train(lsmp_net_handle net, int steps, int batch, double lr, unsigned seed, double weight_c, double *loss)
{
    if (!net || steps < 0 || batch <= 0 || net->nsamp == 0 || !(weight_c >= 0)) return LSMP_E_INVALID;
    const int nblk = (int)cdiv((size_t)batch, TILE);
    int rc = ensure_parts(net, nblk);
    if (rc) return rc;
    for (int s = 0; s < steps; ++s) {
        net->step += 1;
        hipLaunchKernelGGL(k_net_grad, dim3(nblk), dim3(256), 0, net->st, net->P, net->X, net->Y, net->nsamp, batch,
                           seed * 2654435761u + (unsigned)net->step, 0, (float)weight_c, net->part, net->lpart);
        const float bc1 = 1.0f - powf(0.9f, (float)net->step), bc2 = 1.0f - powf(0.999f, (float)net->step);
        hipLaunchKernelGGL(k_net_update, dim3(cdiv(NPARAM, 256)), dim3(256), 0, net->st, net->P, net->m, net->v,
                           net->part, nblk, (float)lr, bc1, bc2, 1, (float *)nullptr, net->lpart, net->loss);
        NET_HIP(hipGetLastError());
    }
    NET_HIP(hipMemcpyAsync(net->hbuf, net->loss, sizeof(double), hipMemcpyDeviceToHost, net->st));
    NET_HIP(hipStreamSynchronize(net->st));
    if (loss) *loss = net->hbuf[0] / batch;
    return LSMP_OK;
}

int lsmp_net_eval_grad(lsmp_net_handle net, const float *x, const float *y, int n, double *loss, float *grad)
{
    if (!net || !x || !y || n <= 0) return LSMP_E_INVALID;
    const size_t nx = (size_t)n * NF;
    float *dX = nullptr;
    float *dY = nullptr;
    NET_HIP(hipMalloc(&dX, nx * sizeof(float)));
    if (hipMalloc(&dY, (size_t)n * sizeof(float)) != hipSuccess) {
        (void)hipFree(dX);
        return LSMP_E_HIP;
    }
    const int nblk = (int)cdiv((size_t)n, TILE);
    int rc = ensure_parts(net, nblk);
    if (!rc && (hipMemcpyAsync(dX, x, nx * sizeof(float), hipMemcpyHostToDevice, net->st) != hipSuccess ||
                hipMemcpyAsync(dY, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, net->st) != hipSuccess))
        rc = LSMP_E_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_net_grad, dim3(nblk), dim3(256), 0, net->st, net->P, dX, dY, (unsigned)n, n, 0u, 1,
                           0.0f, net->part, net->lpart);
        hipLaunchKernelGGL(k_net_update, dim3(cdiv(NPARAM, 256)), dim3(256), 0, net->st, net->P, net->m, net->v,
                           net->part, nblk, 0.0f, 1.0f, 1.0f, 0, net->gsum, net->lpart, net->loss);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(grad, net->gsum, NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->st) != hipSuccess ||
            hipMemcpyAsync(net->hbuf, net->loss, sizeof(double), hipMemcpyDeviceToHost, net->st) != hipSuccess ||
            hipStreamSynchronize(net->st) != hipSuccess)
            rc = LSMP_E_HIP;
        else if (loss)
            *loss = net->hbuf[0] / n;
    }
    (void)hipFree(dX);
    (void)hipFree(dY);
    return rc;
}

int lsmp_reinit_learned(lsmp_handle h, lsmp_net_handle net, double band_cells, int sweeps, double residual_tol,
                        double residual_band, int fallback_mode, double *residual_out, int *fallback, int *n_pseudo)
{
    if (!h || !(band_cells > 0) || sweeps < 1 || !(residual_band > 0)) return LSMP_E_INVALID;
    if (fallback_mode != LSMP_FALLBACK_RESTART && fallback_mode != LSMP_FALLBACK_CONTINUE &&
        fallback_mode != LSMP_FALLBACK_CONTINUE_RELATIVE)
        return LSMP_E_INVALID;
    if (h->dim == 3 && (h->p.dy != h->p.dx || h->p.dz != h->p.dx)) return LSMP_E_INVALID;
    double r = 0.0;
    int fb = 0, np = 0;
    int rc = h->fp32 ? reinit_learned<float>(h, net, band_cells, sweeps, residual_tol, residual_band, fallback_mode,
                                             &r, &fb, &np)
                     : reinit_learned<double>(h, net, band_cells, sweeps, residual_tol, residual_band, fallback_mode,
                                              &r, &fb, &np);
    if (residual_out) *residual_out = r;
    if (fallback) *fallback = fb;
    if (n_pseudo) *n_pseudo = np;
    h->stats.last_redis_iters = np;
    return rc;
}

int lsmp_net_add_samples(lsmp_net_handle net, const float *x, const float *y, int n)
{
    if (!net || !x || !y || n < 0 || (unsigned)n > net->capacity - net->nsamp) return LSMP_E_INVALID;
    if (n == 0) return LSMP_OK;
    NET_HIP(hipMemcpyAsync(net->X + (size_t)net->nsamp * NF, x, (size_t)n * NF * sizeof(float),
                           hipMemcpyHostToDevice, net->st));
    NET_HIP(hipMemcpyAsync(net->Y + net->nsamp, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, net->st));
    NET_HIP(hipStreamSynchronize(net->st));
    net->nsamp += (unsigned)n;
    return LSMP_OK;
}

int lsmp_net_clear_samples(lsmp_net_handle net)
{
    if (!net) return LSMP_E_INVALID;
    net->nsamp = 0;
    return LSMP_OK;
}

int lsmp_net_apply(lsmp_net_handle net, lsmp_handle h, double band_cells, int *n_band)
{
    if (!net || !h || !(band_cells > 0)) return LSMP_E_INVALID;
    if (h->dim == 3 && (h->p.dy != h->p.dx || h->p.dz != h->p.dx)) return LSMP_E_INVALID;
    return h->fp32 ? apply<float>(h, net, band_cells, n_band) : apply<double>(h, net, band_cells, n_band);
}

int lsmp_reinit_residual(lsmp_handle h, double band_cells, double *res)
{
    if (!h || !res || !(band_cells > 0)) return LSMP_E_INVALID;
    return h->fp32 ? residual<float>(h, band_cells, res) : residual<double>(h, band_cells, res);
}

}  // extern "C"
