"""Python host binding of the gfx950 level-set transport library (liblsmp.so).

Mirrors the C ABI of include/lsmp.h one to one; each method names the
reference routine it stands in for.  Arrays cross the boundary as FP64 host
boxes in the reference's storage contract: column-major (i fastest) over
(1-3 : n+3) per direction, 7 z planes in 2D.  In numpy terms a box is a
C-contiguous float64 array of shape (nz+6, ny+6, nx+6) (or (7, ny+6, nx+6)).

There is no CPU fallback: constructing a Solver without the built library or
without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# another build of the library can be named by LSMP_LIBRARY
LIB_PATH = os.environ.get("LSMP_LIBRARY") or os.path.join(HERE, "lib", "liblsmp.so")

FP64, FP32 = 0, 1
FALLBACK_RESTART, FALLBACK_CONTINUE, FALLBACK_CONTINUE_RELATIVE = 0, 1, 2
SCHEME_CRWENO, SCHEME_UCCD = 0, 1
VEL_ROTATION, VEL_VORTEX2D, VEL_LEVEQUE3D = 0, 1, 2

ERRORS = {0: "ok", -1: "invalid parameter or size", -2: "HIP runtime error",
          -3: "non-finite value", -4: "empty interface band", -5: "call out of order"}

# Every entry point declared in include/lsmp.h.
EXPORTS = [
    "lsmp_init", "lsmp_finalize", "lsmp_error_string", "lsmp_upload_phi", "lsmp_download_phi",
    "lsmp_store_phi0", "lsmp_set_velocity", "lsmp_set_velocity_analytic", "lsmp_rk3_advect",
    "lsmp_reinit", "lsmp_mass_vol", "lsmp_mass_fix", "lsmp_interface_l1", "lsmp_get_stats",
    "lsmp_set_profiling", "lsmp_kernel_time", "lsmp_synchronize", "lsmp_set_advection_scheme",
    "lsmp_set_band_lines", "lsmp_band_line_counts", "lsmp_set_narrow_band", "lsmp_component_masses",
    "lsmp_set_marker", "lsmp_get_marker", "lsmp_marker_load", "lsmp_marker_store", "lsmp_combine_markers",
    "lsmp_net_create", "lsmp_net_destroy", "lsmp_net_param_count", "lsmp_net_set_params",
    "lsmp_net_get_params", "lsmp_net_dataset_size", "lsmp_net_collect", "lsmp_net_train",
    "lsmp_net_eval_grad", "lsmp_reinit_learned", "lsmp_reinit_residual", "lsmp_net_add_samples",
    "lsmp_net_clear_samples", "lsmp_net_apply",
    "lsmp_upload_vof", "lsmp_download_vof", "lsmp_store_vof_old", "lsmp_set_face_velocity",
    "lsmp_set_face_velocity_analytic", "lsmp_vof_wlic", "lsmp_clsvof_recon", "lsmp_store_vof0",
    "lsmp_vof_mass_vol", "lsmp_vof_interface_l1", "lsmp_mass_vol_noalloc", "lsmp_energy_momentum",
]
NO_STATUS = ("lsmp_finalize", "lsmp_error_string", "lsmp_net_destroy")
# order of lsmp_energy_momentum's eleven results
ENERGY_KEYS = ("A", "v", "px", "py", "pz", "xc", "yc", "zc", "Ke", "Po", "Q")


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ghc", C.c_int), ("dim", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dt", C.c_double), ("rdt", C.c_double), ("ls_wid", C.c_double),
        ("rho_12", C.c_double), ("inverse", C.c_int), ("precision", C.c_int),
        ("xstart", C.c_double), ("xend", C.c_double), ("ystart", C.c_double),
        ("yend", C.c_double), ("zstart", C.c_double), ("zend", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("last_redis_iters", C.c_int), ("last_mass_iters", C.c_int),
        ("last_red_error", C.c_double), ("t_advect_ms", C.c_double),
        ("t_reinit_ms", C.c_double), ("t_massfix_ms", C.c_double),
        ("n_advect", C.c_longlong), ("n_reinit", C.c_longlong), ("n_massfix", C.c_longlong),
    ]


_DP = C.POINTER(C.c_double)
_lib = None


def load_library(path=LIB_PATH):
    """Load liblsmp.so and declare its signatures.  Raises when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    H = C.c_void_p
    L.lsmp_init.argtypes = [C.POINTER(Params), C.POINTER(H)]
    L.lsmp_finalize.argtypes = [H]
    L.lsmp_finalize.restype = None
    L.lsmp_error_string.argtypes = [C.c_int]
    L.lsmp_error_string.restype = C.c_char_p
    L.lsmp_upload_phi.argtypes = [H, _DP]
    L.lsmp_download_phi.argtypes = [H, _DP]
    L.lsmp_set_marker.argtypes = [H, C.c_int, _DP]
    L.lsmp_get_marker.argtypes = [H, C.c_int, _DP]
    L.lsmp_marker_load.argtypes = [H, C.c_int]
    L.lsmp_marker_store.argtypes = [H, C.c_int]
    L.lsmp_combine_markers.argtypes = [H, C.POINTER(C.c_int)]
    L.lsmp_store_phi0.argtypes = [H]
    L.lsmp_set_velocity.argtypes = [H, _DP, _DP, _DP]
    L.lsmp_set_velocity_analytic.argtypes = [H, C.c_int, C.c_double, C.c_double]
    L.lsmp_rk3_advect.argtypes = [H, C.c_double]
    L.lsmp_reinit.argtypes = [H, C.c_int, C.c_double, _DP, C.POINTER(C.c_int)]
    L.lsmp_mass_vol.argtypes = [H, _DP, _DP]
    L.lsmp_mass_fix.argtypes = [H, C.c_double, _DP, C.POINTER(C.c_int)]
    L.lsmp_interface_l1.argtypes = [H, _DP]
    L.lsmp_energy_momentum.argtypes = [H, C.c_double, C.c_int, _DP]
    L.lsmp_get_stats.argtypes = [H, C.POINTER(Stats)]
    L.lsmp_set_profiling.argtypes = [H, C.c_int]
    L.lsmp_kernel_time.argtypes = [H, C.c_char_p, C.POINTER(C.c_longlong), _DP]
    L.lsmp_synchronize.argtypes = [H]
    L.lsmp_set_advection_scheme.argtypes = [H, C.c_int]
    L.lsmp_set_band_lines.argtypes = [H, C.c_int]
    L.lsmp_band_line_counts.argtypes = [H, C.POINTER(C.c_int)]
    L.lsmp_set_narrow_band.argtypes = [H, C.c_double]
    L.lsmp_component_masses.argtypes = [H, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_int), C.c_int, _DP, C.POINTER(C.c_int)]
    _FP = C.POINTER(C.c_float)
    _IP = C.POINTER(C.c_int)
    L.lsmp_net_create.argtypes = [C.c_int, C.c_uint, C.POINTER(H)]
    L.lsmp_net_destroy.argtypes = [H]
    L.lsmp_net_destroy.restype = None
    L.lsmp_net_param_count.argtypes = []
    L.lsmp_net_set_params.argtypes = [H, _FP, C.c_int]
    L.lsmp_net_get_params.argtypes = [H, _FP, C.c_int]
    L.lsmp_net_dataset_size.argtypes = [H, _IP]
    L.lsmp_net_collect.argtypes = [H, H, C.c_double, C.c_int, _IP]
    L.lsmp_net_train.argtypes = [H, C.c_int, C.c_int, C.c_double, C.c_uint, C.c_double, _DP]
    L.lsmp_net_eval_grad.argtypes = [H, _FP, _FP, C.c_int, _DP, _FP]
    L.lsmp_reinit_learned.argtypes = [H, H, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, _DP, _IP, _IP]
    L.lsmp_reinit_residual.argtypes = [H, C.c_double, _DP]
    L.lsmp_net_add_samples.argtypes = [H, _FP, _FP, C.c_int]
    L.lsmp_net_clear_samples.argtypes = [H]
    L.lsmp_net_apply.argtypes = [H, H, C.c_double, _IP]
    L.lsmp_upload_vof.argtypes = [H, _DP]
    L.lsmp_download_vof.argtypes = [H, _DP]
    L.lsmp_store_vof_old.argtypes = [H]
    L.lsmp_set_face_velocity.argtypes = [H, _DP, _DP, _DP]
    L.lsmp_set_face_velocity_analytic.argtypes = [H, C.c_int, C.c_double, C.c_double]
    L.lsmp_vof_wlic.argtypes = [H, C.c_double]
    L.lsmp_clsvof_recon.argtypes = [H, _IP, _IP]
    L.lsmp_store_vof0.argtypes = [H]
    L.lsmp_vof_mass_vol.argtypes = [H, _DP, _DP]
    L.lsmp_vof_interface_l1.argtypes = [H, _DP]
    L.lsmp_mass_vol_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
    for name in EXPORTS:
        if name not in NO_STATUS:
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


class LsmpError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        msg = load_library().lsmp_error_string(rc).decode()
        raise LsmpError(f"lsmp error {rc}: {msg}")


def _box_ptr(a):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("boxes must be C-contiguous float64 arrays")
    return a.ctypes.data_as(_DP)


def make_params(n, dim=3, dt_cfl=0.1, rdt_cfl=0.1, ls_wid_cells=1.5, rho_12=1.0, inverse=0,
                precision=FP64, ny=None, nz=None):
    """Parameters derived the way the reference derives them from its deck
    (Vortex/lib/data/tree.f90:230-280) on the unit box."""
    ny = n if ny is None else ny
    nz = (n if nz is None else nz) if dim == 3 else 1
    dx = 1.0 / n
    return Params(nx=n, ny=ny, nz=nz, ghc=3, dim=dim, dx=dx, dy=1.0 / ny,
                  dz=(1.0 / nz) if dim == 3 else 1.0, dt=dt_cfl * dx, rdt=rdt_cfl * dx,
                  ls_wid=ls_wid_cells * dx, rho_12=rho_12, inverse=inverse, precision=precision,
                  xstart=0.0, xend=1.0, ystart=0.0, yend=1.0, zstart=0.0, zend=1.0)


class Solver:
    """One device-resident level-set state (the job's phi, nvel and scratch)."""

    def __init__(self, params):
        self.lib = load_library()
        self.params = params
        self.h = C.c_void_p()
        _check(self.lib.lsmp_init(C.byref(params), C.byref(self.h)))
        self.nzb = params.nz + 6 if params.dim == 3 else 7
        self.shape = (self.nzb, params.ny + 6, params.nx + 6)

    def close(self):
        if self.h:
            self.lib.lsmp_finalize(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- data in / out
    def upload_phi(self, box):
        _check(self.lib.lsmp_upload_phi(self.h, _box_ptr(np.ascontiguousarray(box, np.float64))))

    def download_phi(self):
        out = np.empty(self.shape, np.float64)
        _check(self.lib.lsmp_download_phi(self.h, _box_ptr(out)))
        return out

    def store_phi0(self):
        _check(self.lib.lsmp_store_phi0(self.h))

    def set_velocity(self, u, v, w=None):
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        w = np.zeros_like(u) if w is None else np.ascontiguousarray(w, np.float64)
        _check(self.lib.lsmp_set_velocity(self.h, _box_ptr(u), _box_ptr(v), _box_ptr(w)))

    def set_velocity_analytic(self, field, t, t2s):
        _check(self.lib.lsmp_set_velocity_analytic(self.h, field, t, t2s))

    # -- the hot path
    def set_advection_scheme(self, scheme):
        """SCHEME_CRWENO (level_set_rk3_source_weno) or SCHEME_UCCD (source_ccd)."""
        _check(self.lib.lsmp_set_advection_scheme(self.h, int(scheme)))

    def set_band_lines(self, mode):
        """Mass correction: CCD normals on every line (0), on the lines crossing
        the band where that pays (1, the default: 3D grids with more than 4096
        lines in a direction and lines of at most 540 cells) or always (2)."""
        _check(self.lib.lsmp_set_band_lines(self.h, int(mode)))

    def set_narrow_band(self, width_cells):
        """Narrow band of half-width width_cells (3D; 0 turns it off): advection
        and reinitialisation only where |phi| < width_cells * dx."""
        _check(self.lib.lsmp_set_narrow_band(self.h, float(width_cells)))

    def band_line_counts(self):
        """Lines listed by the last mass correction along x, y, z (-1: every line)."""
        c = (C.c_int * 3)()
        _check(self.lib.lsmp_band_line_counts(self.h, c))
        return list(c)

    def component_masses(self, phi_min, query_cells=(), maxcomp=4096, phi_core=None):
        """Connected components of phi > phi_core (default phi_min), grown
        through the cells with phi_min < phi <= phi_core, and their level-set
        masses.  Returns (masses (ncomp,), component of each query cell or -1)."""
        phi_core = phi_min if phi_core is None else phi_core
        q = np.ascontiguousarray(query_cells, dtype=np.int64)
        qc = np.empty(q.size, dtype=np.int32)
        m = np.empty(maxcomp, dtype=np.float64)
        nc = C.c_int()
        _check(self.lib.lsmp_component_masses(self.h, phi_core, phi_min, q.size,
                                              q.ctypes.data_as(C.POINTER(C.c_longlong)),
                                              qc.ctypes.data_as(C.POINTER(C.c_int)), maxcomp,
                                              m.ctypes.data_as(_DP), C.byref(nc)))
        return m[:min(nc.value, maxcomp)].copy(), qc

    def rk3_advect(self, dt=None):
        """level_set_rk3_solver (level_set_rk3_solver.f90:1-94)."""
        _check(self.lib.lsmp_rk3_advect(self.h, self.params.dt if dt is None else dt))

    def reinit(self, btn=1, timestop=-1.0):
        """level_set_rk3_redis(btn[,T]) (level_set_rk3_redis.f90:1-79).
        Returns (red_error, pseudo-steps)."""
        err = C.c_double()
        n = C.c_int()
        _check(self.lib.lsmp_reinit(self.h, btn, timestop, C.byref(err), C.byref(n)))
        return err.value, n.value

    def mass_vol(self):
        """ls_mv, LS part (ns_solver/others.f90:172-213).  Returns (mass, vol)."""
        m = C.c_double()
        v = C.c_double()
        _check(self.lib.lsmp_mass_vol(self.h, C.byref(m), C.byref(v)))
        return m.value, v.value

    def mass_fix(self, imass):
        """mass_preserving_level_set (mass_preserving_level_set.f90:1-74).
        Returns (mass before the last correction, iterations)."""
        m = C.c_double()
        n = C.c_int()
        _check(self.lib.lsmp_mass_fix(self.h, imass, C.byref(m), C.byref(n)))
        return m.value, n.value

    def interface_l1(self):
        """calculate_interface_loss, LS part (output.f90:169-204, 226)."""
        v = C.c_double()
        _check(self.lib.lsmp_interface_l1(self.h, C.byref(v)))
        return v.value

    def energy_momentum(self, mu_12=1.0, vof_marker=False):
        """energy_and_momentum's integrals (output.f90:250-397) before the deck's
        Fr, We, Re and energy_unit scaling: dict of A (surface area), v, px, py,
        pz, xc, yc, zc, Ke, Po, Q.  marker = H(phi), or the VOF (method 3)."""
        out = (C.c_double * len(ENERGY_KEYS))()
        _check(self.lib.lsmp_energy_momentum(self.h, float(mu_12), int(bool(vof_marker)), out))
        return dict(zip(ENERGY_KEYS, list(out)))

    # -- accounting
    def stats(self):
        s = Stats()
        _check(self.lib.lsmp_get_stats(self.h, C.byref(s)))
        return s

    def set_profiling(self, on=True):
        _check(self.lib.lsmp_set_profiling(self.h, 1 if on else 0))

    def kernel_time(self, name):
        n = C.c_longlong()
        ms = C.c_double()
        _check(self.lib.lsmp_kernel_time(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def synchronize(self):
        _check(self.lib.lsmp_synchronize(self.h))

    def reinit_learned(self, net, band_cells=6.0, sweeps=2, residual_tol=1e-2, residual_band=1.5,
                       fallback=FALLBACK_CONTINUE):
        """Learned reinitialisation (no reference counterpart): network on the
        band, `sweeps` plain pseudo-steps, plain fallback (restart from the
        input, or continuation) when the residual over |phi| < residual_band
        dx exceeds residual_tol.  net=None runs the same sequence without the
        network pass (the sweeps-only control).  Returns (residual,
        fallback_taken, n_pseudo)."""
        r = C.c_double()
        fb = C.c_int()
        n = C.c_int()
        _check(self.lib.lsmp_reinit_learned(self.h, net.h if net is not None else None, band_cells, sweeps,
                                            residual_tol, residual_band,
                                            int(fallback), C.byref(r), C.byref(fb), C.byref(n)))
        return r.value, bool(fb.value), n.value

    def reinit_residual(self, band_cells=3.0):
        """Mean | |grad phi| - 1 | over |phi| < band_cells dx."""
        r = C.c_double()
        _check(self.lib.lsmp_reinit_residual(self.h, band_cells, C.byref(r)))
        return r.value

    def mpls_step(self, t, t2s, field, imass, t2red=1, iteration=1, net=None, sweeps=2, residual_tol=1e-2,
                  fallback=FALLBACK_CONTINUE):
        """One MPLS step as the merged branch of mpls runs it
        (interface_solver.f90:123-129) after the driver set the velocity at
        t^{n+1} (main.f90:23-29).  With a trained `net` the reinitialisation
        is the learned one.  Returns (n_pseudo, n_mass)."""
        self.set_velocity_analytic(field, t, t2s)
        self.rk3_advect()
        npseudo = 0
        if iteration % t2red == 0:
            if net is None:
                _, npseudo = self.reinit(1)
            else:
                _, _, npseudo = self.reinit_learned(net, sweeps=sweeps, residual_tol=residual_tol,
                                                    fallback=fallback)
        _, nm = self.mass_fix(imass)
        return npseudo, nm

    # -- CLSVOF (clsvof, interface_solver.f90:133-244)
    def upload_vof(self, box):
        """vof%now from a host box."""
        _check(self.lib.lsmp_upload_vof(self.h, _box_ptr(np.ascontiguousarray(box, np.float64))))

    def download_vof(self):
        out = np.empty(self.shape, np.float64)
        _check(self.lib.lsmp_download_vof(self.h, _box_ptr(out)))
        return out

    def store_vof_old(self):
        """vof%old <- vof%now (the VOF part of p%switch, tree.f90:383-392)."""
        _check(self.lib.lsmp_store_vof_old(self.h))

    def store_vof0(self):
        """vof%tmp2 <- vof%now: the initial VOF of vof_interface_l1."""
        _check(self.lib.lsmp_store_vof0(self.h))

    def vof_mass_vol(self):
        """ls_mv, VOF part (ns_solver/others.f90:215-247).  Returns (mass, vol)."""
        m, v = C.c_double(), C.c_double()
        _check(self.lib.lsmp_vof_mass_vol(self.h, C.byref(m), C.byref(v)))
        return m.value, v.value

    def vof_interface_l1(self):
        """calculate_interface_loss, VOF part: sum |vof0 - vof| / cells."""
        r = C.c_double()
        _check(self.lib.lsmp_vof_interface_l1(self.h, C.byref(r)))
        return r.value

    def set_face_velocity(self, u, v, w=None):
        """vel%x/y/z%old: index i of a box holds the face i + 1/2."""
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        w = np.zeros_like(u) if w is None else np.ascontiguousarray(w, np.float64)
        _check(self.lib.lsmp_set_face_velocity(self.h, _box_ptr(u), _box_ptr(v), _box_ptr(w)))

    def set_face_velocity_analytic(self, field, t, t2s):
        _check(self.lib.lsmp_set_face_velocity_analytic(self.h, field, t, t2s))

    def vof_wlic(self, dt=None):
        """vof_wlic_solver (clsvof.f90:1-273)."""
        _check(self.lib.lsmp_vof_wlic(self.h, self.params.dt if dt is None else dt))

    def clsvof_recon(self):
        """clsvof_recon (clsvof.f90:275-324).  Returns the pseudo-steps of its
        two reinitialisations (btn = 0, then btn = 1)."""
        a, b = C.c_int(), C.c_int()
        _check(self.lib.lsmp_clsvof_recon(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def clsvof_step(self, t, t2s, field, t2red=1, iteration=1, dt=None):
        """One CLSVOF step as the merged branch of clsvof runs it
        (interface_solver.f90:234-244) after the driver set the velocities at
        t^{n+1} and switched vof%old (main.f90:23-27).  Returns the pseudo-steps
        of the reconstruction's reinitialisations ((0, 0) without one)."""
        self.set_velocity_analytic(field, t, t2s)
        self.set_face_velocity_analytic(field, t, t2s)
        self.store_vof_old()
        self.vof_wlic(dt)
        self.rk3_advect(dt)
        if iteration % t2red == 0:
            return self.clsvof_recon()
        return 0, 0

    # -- two-marker MPLS branch (interface_solver.f90:34-121)
    def set_marker(self, mid, box):
        _check(self.lib.lsmp_set_marker(self.h, int(mid), _box_ptr(np.ascontiguousarray(box, np.float64))))

    def get_marker(self, mid):
        out = np.empty(self.shape, np.float64)
        _check(self.lib.lsmp_get_marker(self.h, int(mid), _box_ptr(out)))
        return out

    def marker_load(self, mid):
        """phi <- marker mid (interface_solver.f90:44-56)."""
        _check(self.lib.lsmp_marker_load(self.h, int(mid)))

    def marker_store(self, mid):
        """marker mid <- phi (interface_solver.f90:64-76)."""
        _check(self.lib.lsmp_marker_store(self.h, int(mid)))

    def combine_markers(self):
        """phi <- merge of the two markers (interface_solver.f90:85-111);
        True when some cell is inside both."""
        r = C.c_int()
        _check(self.lib.lsmp_combine_markers(self.h, C.byref(r)))
        return bool(r.value)

    def mpls_step_markers(self, t, t2s, field, imass, marker_imass, merged, t2red=1, iteration=1):
        """One MPLS step as mpls runs it (interface_solver.f90:34-128) after
        the driver set the velocity at t^{n+1}: while not merged, each marker
        is advected, reinitialised and mass corrected against its own initial
        mass, then phi is their merge, reinitialised once when they overlap.
        Returns (merged, n_pseudo, n_mass) with the counts summed over the
        fields transported."""
        self.set_velocity_analytic(field, t, t2s)
        if merged:
            npseudo, nm = 0, 0
            self.rk3_advect()
            if iteration % t2red == 0:
                _, npseudo = self.reinit(1)
            _, nm = self.mass_fix(imass)
            return True, npseudo, nm
        npseudo = nmass = 0
        for mid in (1, 2):
            self.marker_load(mid)
            self.rk3_advect()
            if iteration % t2red == 0:
                npseudo += self.reinit(1)[1]
            nmass += self.mass_fix(marker_imass[mid - 1])[1]
            self.marker_store(mid)
        if self.combine_markers():
            npseudo += self.reinit(1)[1]
            merged = True
        return merged, npseudo, nmass


def mass_vol_noalloc(phi_ptr, nx, ny, nz, dim, dx, dy, dz, ls_wid, rho_12, inverse, out_ptr, stream=None):
    """lsmp_mass_vol_noalloc on device pointers: the (mass, vol) of the FP64
    interior field at phi_ptr into the device doubles at out_ptr, stream
    ordered.  Allocates nothing."""
    _check(load_library().lsmp_mass_vol_noalloc(phi_ptr, nx, ny, nz, dim, dx, dy, dz, ls_wid, rho_12, inverse,
                                                out_ptr, stream))


NF, NH = 128, 64
PARAM_SLICES = {"W1": (0, (NF, NH)), "b1": (NF * NH, (NH,)), "W2": (NF * NH + NH, (NH, NH)),
                "b2": (NF * NH + NH + NH * NH, (NH,)), "w3": (NF * NH + 2 * NH + NH * NH, (NH,)),
                "b3": (NF * NH + 3 * NH + NH * NH, (1,))}


def unpack_params(flat):
    """Split the flat parameter vector into the named arrays (views)."""
    return {k: flat[o:o + int(np.prod(sh))].reshape(sh) for k, (o, sh) in PARAM_SLICES.items()}


class Net:
    """The learned-reinitialisation network and its training set on the GPU."""

    def __init__(self, capacity=0, seed=1):
        self.lib = load_library()
        self.h = C.c_void_p()
        _check(self.lib.lsmp_net_create(int(capacity), int(seed), C.byref(self.h)))
        self.nparam = self.lib.lsmp_net_param_count()

    def close(self):
        if self.h:
            self.lib.lsmp_net_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_params(self):
        out = np.empty(self.nparam, dtype=np.float32)
        _check(self.lib.lsmp_net_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), self.nparam))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {flat.size}")
        _check(self.lib.lsmp_net_set_params(self.h, flat.ctypes.data_as(C.POINTER(C.c_float)), self.nparam))

    def save(self, path):
        np.save(path, self.get_params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    def dataset_size(self):
        n = C.c_int()
        _check(self.lib.lsmp_net_dataset_size(self.h, C.byref(n)))
        return n.value

    def collect(self, solver, band_cells=6.0, stride=1):
        """Plain-reinitialise the solver's phi and store every stride-th band pair."""
        n = C.c_int()
        _check(self.lib.lsmp_net_collect(self.h, solver.h, band_cells, int(stride), C.byref(n)))
        return n.value

    def add_samples(self, x, y):
        """Append host pairs x (n, 128), y (n,) to the training set."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != NF or y.shape != (x.shape[0],):
            raise ValueError("x must be (n, 128) and y (n,)")
        _check(self.lib.lsmp_net_add_samples(self.h, x.ctypes.data_as(C.POINTER(C.c_float)),
                                             y.ctypes.data_as(C.POINTER(C.c_float)), x.shape[0]))

    def clear_samples(self):
        _check(self.lib.lsmp_net_clear_samples(self.h))

    def apply(self, solver, band_cells=6.0):
        """The network pass alone on the solver's phi; returns the band size."""
        n = C.c_int()
        _check(self.lib.lsmp_net_apply(self.h, solver.h, band_cells, C.byref(n)))
        return n.value

    def train(self, steps, batch=4096, lr=1e-3, seed=0, weight_c=0.0):
        """Adam steps; weight_c > 0 weights each pair by c^2 / (y^2 + c^2)."""
        loss = C.c_double()
        _check(self.lib.lsmp_net_train(self.h, int(steps), int(batch), float(lr), int(seed), float(weight_c),
                                       C.byref(loss)))
        return loss.value

    def eval_grad(self, x, y):
        """Loss and gradient on host samples x (n, 128), y (n,)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != NF or y.shape != (x.shape[0],):
            raise ValueError("x must be (n, 128) and y (n,)")
        g = np.empty(self.nparam, dtype=np.float32)
        loss = C.c_double()
        _check(self.lib.lsmp_net_eval_grad(self.h, x.ctypes.data_as(C.POINTER(C.c_float)),
                                           y.ctypes.data_as(C.POINTER(C.c_float)), x.shape[0], C.byref(loss),
                                           g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g
