"""ctypes access to the test oracle (oracle/lsmp_oracle.c) and to the
reference's own Fortran units (oracle/_ref/libref_pieces.so).

Test infrastructure only: the oracle is the checker, never the thing measured
or shipped.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# LSMP_ORACLE_SO names another build of the same source (the OpenMP one of the
# CPU baseline)
ORACLE_SO = os.environ.get("LSMP_ORACLE_SO") or os.path.join(ROOT, "oracle", "_build", "liblsmp_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_pieces.so")

DP = C.POINTER(C.c_double)

FIELDS = [
    "PHI", "PHI_TMP", "PHI0", "U", "V", "W",
    "HEAVY", "DELTA", "SIGN", "SIGN_TMP", "GRAD",
    "SX1", "SX2", "SY1", "SY2", "SZ1", "SZ2",
    "SSX1", "SSX2", "SSY1", "SSY2", "SSZ1", "SSZ2",
    "S3", "L1", "L2", "L3",
    "NX", "NY", "NZ", "NXX", "NYY", "NZZ",
    "VOF",
    "MK1", "MK2",
    "VOF_OLD", "UF", "VF", "WF", "VOF_TMP",
]
F = {n: i for i, n in enumerate(FIELDS)}


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dim", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dt", C.c_double), ("rdt", C.c_double), ("ls_wid", C.c_double),
        ("rho_12", C.c_double), ("inverse", C.c_int),
        ("xstart", C.c_double), ("xend", C.c_double),
        ("ystart", C.c_double), ("yend", C.c_double),
        ("zstart", C.c_double), ("zend", C.c_double),
    ]


# order of the energy_and_momentum integrals (oracle and library alike)
ENERGY_KEYS = ("A", "v", "px", "py", "pz", "xc", "yc", "zc", "Ke", "Po", "Q")


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_field.restype = DP
        L.orc_field.argtypes = [C.c_void_p, C.c_int]
        L.orc_box_size.restype = C.c_long
        L.orc_box_size.argtypes = [C.c_void_p]
        L.orc_get.restype = C.c_double
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_set.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.orc_solve_tridiagonal.argtypes = [DP, DP, DP, DP, DP, C.c_int]
        L.orc_wenojs_flux_split.argtypes = [DP, DP, DP, DP, C.c_int]
        L.orc_crweno_flux_split.argtypes = [DP, DP, DP, DP, C.c_int]
        L.orc_phyn.restype = C.c_double
        L.orc_phyn.argtypes = [C.c_double] * 4
        L.orc_ccd_line.argtypes = [C.c_int, C.c_double, DP, DP, DP]
        L.orc_uccd_line.argtypes = [C.c_int, C.c_double, DP, DP, DP]
        for name in ["orc_ls_funs", "orc_ls_mv", "orc_rk3_source", "orc_rk3_advect", "orc_rk3_source_uccd",
                     "orc_redis_gradient", "orc_surface_norms", "orc_store_initial"]:
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_bc.argtypes = [C.c_void_p, DP]
        L.orc_redis_lambda.argtypes = [C.c_void_p, C.c_int]
        L.orc_redis.restype = C.c_int
        L.orc_redis.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_mass_fix.restype = C.c_int
        L.orc_mass_fix.argtypes = [C.c_void_p]
        L.orc_interface_l1.restype = C.c_double
        L.orc_interface_l1.argtypes = [C.c_void_p]
        L.orc_velocity.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_initial_phi.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_mpls_step.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.orc_combine_markers.argtypes = [C.c_void_p]
        L.orc_mpls_step_markers.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.orc_velocity_faces.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_vof_wlic.argtypes = [C.c_void_p]
        L.orc_clsvof_recon.argtypes = [C.c_void_p]
        L.orc_clsvof_step.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.orc_energy_momentum.argtypes = [C.c_void_p, C.c_double, C.c_int, DP]
        _lib = L
    return _lib


_ref = None


def ref():
    """The reference's TDMA, CCD, CRWENO5 and WENO-HJ units, or None when
    they were not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        L = C.CDLL(REF_SO)
        L.ref_solve_tridiagonal.argtypes = [DP, DP, DP, DP, DP, C.c_int]
        L.ref_ccd_line.argtypes = [C.c_int, C.c_double, DP, DP, DP]
        L.ref_uccd_line.argtypes = [C.c_int, C.c_double, DP, DP, DP]
        L.ref_ccd_tables.argtypes = [C.c_int, C.c_double, DP]
        L.ref_crweno_flux_split.argtypes = [DP, DP, DP, DP, C.c_int]
        L.ref_phyn.argtypes = [C.c_double] * 4
        L.ref_phyn.restype = C.c_double
        _ref = L
    return _ref


def make_params(n, dim=3, dt_cfl=0.1, rdt_cfl=0.1, ls_wid_cells=1.5, rho_12=1.0,
                inverse=0, ny=None, nz=None):
    """Parameters the way the reference derives them from its deck
    (Vortex/lib/data/tree.f90:230-280): n cells per unit length on the unit
    box, dx = 1/n, dt = deck*dx, rdt = deck*dx, ls_wid = deck*dx."""
    ny = n if ny is None else ny
    nz = (n if nz is None else nz) if dim == 3 else 1
    dx = 1.0 / n
    dy = 1.0 / ny
    dz = 1.0 / nz if dim == 3 else 1.0
    return OrcParams(nx=n, ny=ny, nz=nz, dim=dim, dx=dx, dy=dy, dz=dz,
                     dt=dt_cfl * dx, rdt=rdt_cfl * dx, ls_wid=ls_wid_cells * dx,
                     rho_12=rho_12, inverse=inverse,
                     xstart=0.0, xend=1.0, ystart=0.0, yend=1.0, zstart=0.0, zend=1.0)


class Oracle:
    """Owns one oracle state; fields are exposed as numpy views shaped
    (nzb, nyb, nxb) of the column-major box (i fastest)."""

    def __init__(self, params):
        self.p = params
        self.h = lib().orc_create(C.byref(params))
        self.nxb = params.nx + 6
        self.nyb = params.ny + 6
        self.nzb = (params.nz + 6) if params.dim == 3 else 7

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def field(self, name):
        ptr = lib().orc_field(self.h, F[name])
        n = self.nxb * self.nyb * self.nzb
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.nzb, self.nyb, self.nxb)

    def interior(self, name):
        a = self.field(name)
        if self.p.dim == 3:
            return a[3:-3, 3:-3, 3:-3]
        return a[3:4, 3:-3, 3:-3]

    def get(self, name):
        return lib().orc_get(self.h, name.encode())

    def set(self, name, v):
        lib().orc_set(self.h, name.encode(), float(v))

    def __getattr__(self, name):
        fn = getattr(lib(), "orc_" + name)

        def call(*args):
            return fn(self.h, *args)
        return call

    def bc_field(self, name):
        lib().orc_bc(self.h, lib().orc_field(self.h, F[name]))

    def energy_momentum(self, mu_12=1.0, vof_marker=False):
        """energy_and_momentum's integrals (output.f90:250-397) before the deck's
        scaling: dict of A, v, px, py, pz, xc, yc, zc, Ke, Po, Q."""
        out = np.zeros(11)
        lib().orc_energy_momentum(self.h, float(mu_12), int(bool(vof_marker)), _dp(out))
        return dict(zip(ENERGY_KEYS, out.tolist()))
