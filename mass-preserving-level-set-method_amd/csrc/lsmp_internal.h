// lsmp_internal.h -- tile constants and kernel launchers shared by the kernel
// and host translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace lsmp {

constexpr int CSEG = 16;   // rows per CCD checkpoint segment

// every line of a direction goes through the partitioned (one wave or 32
// lanes per line) solves up to these counts
constexpr int PART_MAX_LINES_X = 32768, PART_MAX_LINES_STRIDED = 4096;
constexpr int KCC = 17;   // CCD rows per lane at most (9 up to 284 cells, 17 up to 540 with 32 lanes per line, 1084 with 64)

// listed band lines solved with 32 lanes per line up to these counts (x lines
// are contiguous; y and z lines are strided, and 32 lanes on one of them read
// 32 separate cache lines per row)
#ifndef LSMP_PART_LISTED_X
#define LSMP_PART_LISTED_X (1 << 30)
#endif
#ifndef LSMP_PART_LISTED_STRIDED
#define LSMP_PART_LISTED_STRIDED (1 << 30)
#endif
constexpr int PART_MAX_LISTED_X = LSMP_PART_LISTED_X, PART_MAX_LISTED_STRIDED = LSMP_PART_LISTED_STRIDED;

// reinitialisation tile: RTX x RTY interior cells, one ring for the lambda
// box, four for the WENO-HJ stencil; RKC planes per workgroup in z
constexpr int RTX = 32, RTY = 16;
constexpr int REX = RTX + 2, REY = RTY + 2;
constexpr int RPX = RTX + 8, RPY = RTY + 8;
constexpr int RNT = 640;
static_assert(REX * REY <= RNT, "reinit tile needs one thread per ring cell");

// 3D reinitialisation tile: R3X x R3Y interior cells, one thread each (a
// 4-wave workgroup); RKC planes per workgroup in z.
#ifndef LSMP_R3Y
#define LSMP_R3Y 8
#endif
constexpr int R3X = 32, R3Y = LSMP_R3Y;
constexpr int P3X = R3X + 8, P3Y = R3Y + 8;
constexpr int RKC = 32;
// the narrow band's workgroups march RKC_NB planes: a workgroup without a tube
// cell is skipped, and an active one's march is the kernel's critical path
constexpr int RKC_NB = 8;

// The pseudo-time loop's control words (k_redis_ctl): ctl[0..3], then the
// running max |dphi| striped over CTL_SLOTS words by workgroup, so that the
// workgroups of a stage, which end close together, do not queue their
// atomics on one address.
constexpr int CTL_SLOTS = 64;
constexpr int CTL_WORDS = 4 + CTL_SLOTS;
__device__ inline unsigned long long *ctl_slot(unsigned long long *ctl, unsigned b)
{
    return ctl + 4 + b % (unsigned)CTL_SLOTS;
}

struct RedisArgs {
    int nx, ny, nz;
    double inv12dx, inv12dy, inv12dz, invdx, invdy, invdz, ls_wid, inv_ls_wid, rdt;
    // 3.0 as an argument, so that it sits in a scalar register: a three-operand
    // VALU instruction of gfx9 takes no literal operand
    double three = 3.0;
};

struct CcdArgs {
    double dx, dy, dz, dv, ls_wid, rho_12;
    int inverse;
};

struct VelArgs {
    int nx, ny, nz, field;
    double xstart, ystart, zstart, dx, dy, dz, ct, omega;
};

// rows per lane at most in the partitioned line solves (8 up to 257 cells,
// 16 up to 513)
constexpr int PKX = 16;
// lines of an advection pass from which the tiled solve (lsmp_adv_tile.hip)
// takes over from the paired, twisted and partitioned ones
constexpr int TILE_MIN_LINES = 4096;

// Narrow band: the RK stage of the last direction updates a cell only when its
// x line and, in 3D, its y line cross the tube (fx, fy: the band line flags of
// those directions); fx == nullptr updates every cell.
struct TubeMask {
    const unsigned char *fx = nullptr, *fy = nullptr;
    int nx = 0, ny = 0;
    __device__ bool in(size_t a) const
    {
        if (!fx) return true;
        const size_t i = a % (size_t)nx, r = a / (size_t)nx, j = r % (size_t)ny, k = r / (size_t)ny;
        return fx[j + (size_t)ny * k] && (!fy || fy[i + (size_t)nx * k]);
    }
};

// listed strided lines solved one wave per line up to this count
#ifndef LSMP_PART_LISTED_ADV
#define LSMP_PART_LISTED_ADV 16384
#endif
constexpr int PART_MAX_LISTED_ADV = LSMP_PART_LISTED_ADV;

// lines == nullptr: every line; else the nlist listed lines (the narrow band)
template <typename T>
void launch_adv_pass(hipStream_t st, int dir, int mode, int stage, T *phi, const T *vel, T *R, T *L1, T *L2,
                     T *scr, int nx, int ny, int nz, T dd, T dt, const int *lines = nullptr, int nlist = 0,
                     const TubeMask &tm = TubeMask());
// narrow band: flags of the 3D reinitialisation workgroups with a cell |phi| < wband
template <typename T>
void launch_tube_tiles(hipStream_t st, const T *phi, int nx, int ny, int nz, double wband, unsigned char *tiles);
// 3D reinitialisation stage (lsmp_redis.hip): band = false launches the pass
// over every cell with lambda = 0, band = true (btn = 1 only) the pass over the
// listed band cells with lambda.  gr holds |grad| of the band cells, list
// their cell indices (room for every cell); count is the list's length for
// this stage (zero when the stage starts) and next the one for the next
// stage, which the band pass clears.
template <typename T> struct RedisBand {
    T *gr = nullptr;
    unsigned *list = nullptr, *count = nullptr, *next = nullptr;
};
template <typename T>
void launch_adv_tile(hipStream_t st, int mode, int stage, T *phi, const T *vel, T *R, T *L1, T *L2, int nlines, int n,
                     int lw, long ostride, long istride, long estride, T dd, T dt, const int *lines,
                     const TubeMask &tm);
template <typename T>
void launch_redis3d(hipStream_t st, int stage, int btn, const T *phi, T *phi_out, const T *phi0, const T *S0,
                    const RedisArgs &ra, unsigned long long *ctl, const unsigned char *tiles, const RedisBand<T> &bd,
                    bool band);
template <typename T>
void launch_combine_markers(hipStream_t st, const T *m1, const T *m2, T *phi, size_t n, unsigned *flag);
template <typename T>
void launch_redis_stage2d(hipStream_t st, int stage, int btn, const T *phi, T *phi_out, const T *phi0, const T *S0,
                          const RedisArgs &ra, unsigned long long *ctl);
template <typename T> void launch_sign_init(hipStream_t st, const T *phi, T *S0, size_t n, T ls_wid, int raw);
template <typename T>
void launch_grad_max(hipStream_t st, int dim, const T *phi, const T *S0, const RedisArgs &ra,
                     unsigned long long *gmax);
template <typename T> void launch_scale_by_max(hipStream_t st, T *phi, size_t n, const unsigned long long *gmax);
template <typename T>
void launch_ccd_pass(hipStream_t st, int dir, int dim, bool final_pass, const T *phi, const T *coef, T *scr, T *out,
                     const T *dxv, const T *dyv, T *gr, double *partial, int nx, int ny, int nz, const CcdArgs &ca,
                     int *nblocks_out, const int *lines = nullptr, int nlist = 0);
// mass and volume partial sums of the lines of direction dir with lflag 0,
// each line in up to LINE_SEGMENTS segments (one block of partial sums per 256
// lines and segment)
constexpr int LINE_SEGMENTS = 8;
template <typename T>
void launch_line_hsums(hipStream_t st, int dir, const T *phi, const unsigned char *lflag, int nx, int ny, int nz,
                       const CcdArgs &ca, double *partial, int *nblocks_out);
// the mass update on the cells of the nlist listed lines of direction dir
template <typename T>
void launch_mass_update_lines(hipStream_t st, int dir, T *phi, const T *gr, T ls_wid, const double *scal,
                              const int *lines, int nlist, int nx, int ny, int nz);

// Lines that cross the interface band: one flag byte per line of each
// direction (x: ny nz, y: nx nz, z: nx ny), each direction's flags starting
// on a 16-byte boundary, and the ascending lists of the flagged lines.
constexpr size_t pad16(size_t n) { return (n + 15) & ~(size_t)15; }
__host__ __device__ constexpr size_t band_flag_offset(int d, size_t nfx, size_t nfy)
{
    return d == 0 ? 0 : d == 1 ? pad16(nfx) : pad16(nfx) + pad16(nfy);
}
constexpr size_t band_flag_bytes(size_t nfx, size_t nfy, size_t nfz) { return pad16(nfx) + pad16(nfy) + pad16(nfz); }
// flags of the lines with a cell |phi| < wband, their lists (listx, listy and,
// in 3D, listz) and counts (count[0 .. dim-1]); clamp (the narrow band's tube,
// 3D): phi is first clamped to [-wband, wband] and the lines of the band cells'
// six neighbours are flagged as well
template <typename T>
void launch_band_lines(hipStream_t st, int dim, T *phi, int nx, int ny, int nz, double wband,
                       unsigned char *flags, unsigned *count, int *listx, int *listy, int *listz, bool clamp);
void launch_finalize3(hipStream_t st, const double *partial, int nblk, double imass, double *scal);
// zero n bytes; the ordered lists and counts of the flagged lines (flags laid
// out as band_flag_offset; one workgroup per direction)
void launch_zero_bytes(hipStream_t st, unsigned char *p, size_t n);
void launch_compact_lines(hipStream_t st, int dim, const unsigned char *flags, int nfx, int nfy, int nfz,
                          int *listx, int *listy, int *listz, unsigned *count);
template <typename T>
void launch_mass_update(hipStream_t st, T *phi, const T *gr, size_t n, T ls_wid, const double *scal);
template <typename T>
void launch_mass_vol(hipStream_t st, const T *phi, size_t n, const CcdArgs &ca, double *partial, int nblocks);
template <typename T>
void launch_l1(hipStream_t st, const T *phi, const T *phi0, size_t n, double ls_wid, double *partial, int nblocks);
template <typename T> void launch_velocity(hipStream_t st, T *u, T *v, T *w, const VelArgs &va, double *tab);
template <typename T>
void launch_uccd_pass(hipStream_t st, int dir, int mode, int stage, const T *phi, T *phi_out, const T *vel, T *R,
                      T *L1, T *L2, const T *coef, T *ck, int nx, int ny, int nz, T dx, T dt);
template <typename TO, typename TI> void launch_convert(hipStream_t st, TO *out, const TI *in, size_t n);

// CLSVOF (lsmp_clsvof.hip).  Face velocities of a prescribed field at time
// va.ct / va.omega into uf, vf, wf; tab holds 3 (nx + ny + nz) doubles.
struct FaceVelArgs;
template <typename T>
void launch_face_velocity(hipStream_t st, T *uf, T *vf, T *wf, const FaceVelArgs &va, double xend, double yend,
                          double zend, double *tab);
// one direction's WLIC-THINC pass, in -> out (n2 null in 2D; clamp on the
// last direction)
template <typename T>
void launch_wlic_pass(hipStream_t st, int dir, bool clamp, const T *in, T *out, const T *old, const T *vel,
                      const T *n0, const T *n1, const T *n2, int nx, int ny, int nz, double dt, double dd);
// flags and ordered lists of the lines (x, y and, in 3D, z) through the cells
// that a WLIC step can find cut (wlic_may_cut)
template <typename T>
void launch_cut_lines(hipStream_t st, int dim, const T *vof, int nx, int ny, int nz, unsigned char *flags,
                      unsigned *count, int *listx, int *listy, int *listz);
template <typename T> void launch_recon_start(hipStream_t st, T *phi, const T *vof, T *keep, size_t n);
template <typename T> void launch_recon_keep(hipStream_t st, T *phi, const T *keep, size_t n, double w);
// block partial sums (3 per block) over the VOF: volume and mass (l1 == 0) or
// sum |vof0 - vof| (l1 != 0)
template <typename T>
void launch_vof_sums(hipStream_t st, const T *vof, const T *vof0, size_t n, double dv, double r12, int inverse,
                     int l1, double *partial, int nblocks);
void launch_zero_u64(hipStream_t st, unsigned long long *p, int n);

// energy_and_momentum (lsmp_energy.hip): eleven block sums per workgroup
// over the interior cells and their fixed-order reduction into out[11]
constexpr int ENERGY_SUMS = 11;
struct EnergyArgs {
    int nx, ny, nz, inverse;
    double dx, dv, ls_wid, rho_12, mu_12;
    double xstart, ystart, zstart, dxc, dyc, dzc;   // cell-centre coordinates
};
template <typename T>
void launch_energy_sums(hipStream_t st, int dim, bool vof_marker, const T *phi, const T *vof, const T *u, const T *v,
                        const T *w, const T *ten, const EnergyArgs &ea, double *partial, int nblocks);
void launch_energy_finalize(hipStream_t st, const double *partial, int nblk, double dx, double *out);
void launch_redis_ctl(hipStream_t st, unsigned long long *ctl, int op);

}  // namespace lsmp
