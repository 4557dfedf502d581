// lsmp_metrics.hip -- per-interface mass: connected components of the region
// phi > phi_core (6-connected, interior cells), each grown through the cells
// with phi_min < phi <= phi_core (the smeared half of the interface band),
// and the level-set mass of each.
//
// No counterpart in the reference, which conserves and reports the total mass
// only (ls_mv, ns_solver/others.f90:172-213).  The per-component integrand is
// ls_mv's: rho(H) H dV with H from ls_funs (others.f90:1-45), in FP64.
//
// Labelling: every cell of the region starts with its own index; repeated
// min-propagation over the 6 neighbours (each sweep also follows the label
// pointer once, which halves the remaining distance) until a sweep changes
// nothing.  The component's label is then the smallest cell index in it.
// Growth: one face-neighbour layer per sweep (read one label array, write the
// other), so a band cell joins the component it is fewest steps from, the
// lower label on ties; band cells no component reaches belong to none.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "lsmp_ctx.h"
#include "lsmp_device.h"

namespace lsmp {

constexpr int NOLAB = 0x7fffffff;

template <typename T>
__global__ void k_cc_init(const T *__restrict__ phi, size_t n, T thr, int *__restrict__ lab)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) lab[q] = phi[q] > thr ? (int)q : NOLAB;
}

__global__ void k_cc_sweep(int *__restrict__ lab, int nx, int ny, int nz, unsigned *changed)
{
    const size_t n = (size_t)nx * ny * nz;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int l = lab[q];
    if (l == NOLAB) return;
    const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
    int m = l;
    auto nb = [&](int a, int b, int c) {
        const int v = lab[((size_t)c * ny + b) * nx + a];
        if (v < m) m = v;
    };
    if (i > 0) nb(i - 1, j, k);
    if (i < nx - 1) nb(i + 1, j, k);
    if (j > 0) nb(i, j - 1, k);
    if (j < ny - 1) nb(i, j + 1, k);
    if (k > 0) nb(i, j, k - 1);
    if (k < nz - 1) nb(i, j, k + 1);
    // pointer jump: the label of the cell this label names
    const int mm = lab[m];
    if (mm < m) m = mm;
    if (m < l) {
        atomicMin(&lab[q], m);
        *changed = 1u;
    }
}

template <typename T>
__global__ void k_cc_grow(const T *__restrict__ phi, const int *__restrict__ lab, int *__restrict__ lab_out,
                          int nx, int ny, int nz, T phi_min, unsigned *changed)
{
    const size_t n = (size_t)nx * ny * nz;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int l = lab[q];
    if (l != NOLAB || !(phi[q] > phi_min)) {
        lab_out[q] = l;
        return;
    }
    const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
    int m = NOLAB;
    auto nb = [&](int a, int b, int c) {
        const int v = lab[((size_t)c * ny + b) * nx + a];
        if (v < m) m = v;
    };
    if (i > 0) nb(i - 1, j, k);
    if (i < nx - 1) nb(i + 1, j, k);
    if (j > 0) nb(i, j - 1, k);
    if (j < ny - 1) nb(i, j + 1, k);
    if (k > 0) nb(i, j, k - 1);
    if (k < nz - 1) nb(i, j, k + 1);
    lab_out[q] = m;
    if (m != NOLAB) *changed = 1u;
}

// roots (lab[q] == q) get consecutive component ids in index order
__global__ void k_cc_roots(const int *__restrict__ lab, size_t n, int *__restrict__ rootflag)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) rootflag[q] = lab[q] == (int)q ? 1 : 0;
}

template <typename T>
__global__ void k_cc_mass(const T *__restrict__ phi, const int *__restrict__ lab, const int *__restrict__ cid,
                          size_t n, double ls_wid, double rho_12, int inverse, double dv, double *__restrict__ mass)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int l = lab[q];
    if (l == NOLAB) return;
    double h, d;
    ls_funs<double>((double)phi[q], ls_wid, h, d);
    const double rho = inverse ? rho_12 * h + (1.0 - h) : h + rho_12 * (1.0 - h);
    atomicAdd(&mass[cid[l]], rho * h * dv);
}

}  // namespace lsmp

using namespace lsmp;

namespace {

thread_local char g_merr[256];

#define MET_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            snprintf(g_merr, sizeof g_merr, "%s: %s", #call, hipGetErrorString(e_));               \
            fprintf(stderr, "lsmp: %s\n", g_merr);                                                 \
            return LSMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

unsigned cdiv(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

template <typename T>
int components(lsmp_ctx *c, double phi_core, double phi_min, int nquery, const long long *query, int *query_comp,
               int maxcomp, double *comp_mass, int *ncomp)
{
    const size_t n = c->ncell;
    int *lab = nullptr, *cid = nullptr;
    unsigned *flag = nullptr;
    double *mass = nullptr;
    int rc = LSMP_OK;
    std::vector<int> hcid;
    auto cleanup = [&]() {
        if (lab) (void)hipFree(lab);
        if (cid) (void)hipFree(cid);
        if (flag) (void)hipFree(flag);
        if (mass) (void)hipFree(mass);
    };
#define MET_TRY(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            fprintf(stderr, "lsmp: %s: %s\n", #call, hipGetErrorString(e_));                       \
            cleanup();                                                                             \
            return LSMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)
    MET_TRY(hipMalloc(&lab, n * sizeof(int)));
    MET_TRY(hipMalloc(&cid, n * sizeof(int)));
    MET_TRY(hipMalloc(&flag, sizeof(unsigned)));
    hipLaunchKernelGGL((k_cc_init<T>), dim3(cdiv(n, 256)), dim3(256), 0, c->st, (const T *)c->phi, n,
                       (T)phi_core, lab);
    MET_TRY(hipGetLastError());
    // propagate until a sweep changes nothing (checked every 8 sweeps)
    unsigned hflag = 1;
    for (int sweep = 0; hflag && sweep < 1000000; sweep += 8) {
        MET_TRY(hipMemsetAsync(flag, 0, sizeof(unsigned), c->st));
        for (int q = 0; q < 8; ++q)
            hipLaunchKernelGGL(k_cc_sweep, dim3(cdiv(n, 256)), dim3(256), 0, c->st, lab, c->nx, c->ny, c->nz, flag);
        MET_TRY(hipGetLastError());
        MET_TRY(hipMemcpyAsync(&hflag, flag, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        MET_TRY(hipStreamSynchronize(c->st));
    }
    if (phi_min < phi_core) {
        // grow the components through the band, one layer per sweep; cid is
        // the second label array until the ids are numbered
        hflag = 1;
        int *a = lab, *b = cid;
        for (int sweep = 0; hflag && sweep < 1000000; ++sweep) {
            MET_TRY(hipMemsetAsync(flag, 0, sizeof(unsigned), c->st));
            hipLaunchKernelGGL((k_cc_grow<T>), dim3(cdiv(n, 256)), dim3(256), 0, c->st, (const T *)c->phi, a, b,
                               c->nx, c->ny, c->nz, (T)phi_min, flag);
            MET_TRY(hipGetLastError());
            MET_TRY(hipMemcpyAsync(&hflag, flag, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
            MET_TRY(hipStreamSynchronize(c->st));
            std::swap(a, b);
        }
        if (a != lab) MET_TRY(hipMemcpyAsync(lab, a, n * sizeof(int), hipMemcpyDeviceToDevice, c->st));
    }
    // component ids: prefix numbering of the roots, on the host (labels are
    // cell indices, so a host pass over the root flags is exact and ordered)
    hipLaunchKernelGGL(k_cc_roots, dim3(cdiv(n, 256)), dim3(256), 0, c->st, lab, n, cid);
    MET_TRY(hipGetLastError());
    hcid.resize(n);
    MET_TRY(hipMemcpyAsync(hcid.data(), cid, n * sizeof(int), hipMemcpyDeviceToHost, c->st));
    MET_TRY(hipStreamSynchronize(c->st));
    int nc = 0;
    for (size_t q = 0; q < n; ++q) hcid[q] = hcid[q] ? nc++ : -1;
    if (ncomp) *ncomp = nc;
    MET_TRY(hipMemcpyAsync(cid, hcid.data(), n * sizeof(int), hipMemcpyHostToDevice, c->st));
    MET_TRY(hipMalloc(&mass, (size_t)(nc > 0 ? nc : 1) * sizeof(double)));
    MET_TRY(hipMemsetAsync(mass, 0, (size_t)(nc > 0 ? nc : 1) * sizeof(double), c->st));
    const double dv = c->p.dx * c->p.dy * c->p.dz;
    hipLaunchKernelGGL((k_cc_mass<T>), dim3(cdiv(n, 256)), dim3(256), 0, c->st, (const T *)c->phi, lab, cid, n,
                       c->p.ls_wid, c->p.rho_12, c->p.inverse, dv, mass);
    MET_TRY(hipGetLastError());
    std::vector<double> hm((size_t)(nc > 0 ? nc : 1));
    MET_TRY(hipMemcpyAsync(hm.data(), mass, hm.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    std::vector<int> hl;
    if (nquery > 0) {
        hl.resize(n);
        MET_TRY(hipMemcpyAsync(hl.data(), lab, n * sizeof(int), hipMemcpyDeviceToHost, c->st));
    }
    MET_TRY(hipStreamSynchronize(c->st));
    for (int q = 0; q < nquery; ++q) {
        const long long cell = query[q];
        int comp = -1;
        if (cell >= 0 && (size_t)cell < n && hl[cell] != NOLAB) comp = hcid[hl[cell]];
        query_comp[q] = comp;
    }
    for (int q = 0; q < nc && q < maxcomp; ++q) comp_mass[q] = hm[q];
    cleanup();
    return rc;
#undef MET_TRY
}

}  // namespace

extern "C" int lsmp_component_masses(lsmp_handle h, double phi_core, double phi_min, int nquery,
                                     const long long *query_cells, int *query_comp, int maxcomp, double *comp_mass,
                                     int *ncomp)
{
    if (!h || nquery < 0 || maxcomp < 0 || (nquery > 0 && (!query_cells || !query_comp)) ||
        (maxcomp > 0 && !comp_mass) || !(phi_min <= phi_core))
        return LSMP_E_INVALID;
    return h->fp32 ? components<float>(h, phi_core, phi_min, nquery, query_cells, query_comp, maxcomp, comp_mass,
                                       ncomp)
                   : components<double>(h, phi_core, phi_min, nquery, query_cells, query_comp, maxcomp, comp_mass,
                                        ncomp);
}
