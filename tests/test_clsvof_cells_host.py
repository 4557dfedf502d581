"""The CLSVOF kernels' per-cell code, built for the host, against the oracle.

k_face_velocity and k_wlic_pass are thin loops over the functions of the
library's lsmp_clsvof.h (the line of a cell, the WLIC-THINC face flux, the
cell update, the face-velocity tables).  tests/host_cells compiles that header
with the oracle's flags and runs the same loops cell by cell; here they must
reproduce the oracle bit for bit: the face velocities of the three prescribed
fields, and vof_wlic_solver's three passes (fluxes, divergence term, clamp,
index clamping at the walls) on a fresh sphere or circle and after a few
CLSVOF steps, through both cell updates (the one that loads its stencil up
front as well).  The kernels add only the thread-to-cell map and the store.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
CELLS = os.path.join(HERE, "host_cells")
LIB = os.path.join(CELLS, "_build", "libclsvof_cells.so")
DP = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def cells():
    subprocess.run(["make", "-s", "-C", CELLS], check=True)
    L = C.CDLL(LIB)
    L.hc_face_velocity.argtypes = [C.c_int] * 4 + [C.c_double] * 11 + [DP] * 3
    L.hc_vof_wlic.argtypes = [C.c_int] * 4 + [C.c_double] * 4 + [DP] * 8 + [C.c_int]
    L.hc_vof_wlic_n.argtypes = [C.c_int] * 4 + [C.c_double] * 4 + [DP] * 8 + [C.c_int] * 2
    BP = C.POINTER(C.c_ubyte)
    L.hc_cut_line_flags.argtypes = [DP, C.c_int, C.c_int, C.c_int, BP, BP, BP]
    return L


def ptr(a):
    return a.ctypes.data_as(DP)


def interior(o, name):
    return np.ascontiguousarray(o.interior(name))


def face_velocity(cells, o, field, t, t2s):
    p = o.p
    out = [np.empty((p.nz, p.ny, p.nx)) for _ in range(3)]
    ct = np.cos(np.pi * t / t2s)
    omega = 2.0 * np.pi / t2s
    cells.hc_face_velocity(p.nx, p.ny, p.nz, field, p.xstart, p.xend, p.ystart, p.yend, p.zstart, p.zend,
                           p.dx, p.dy, p.dz, ct, omega, *(ptr(a) for a in out))
    return out


@pytest.mark.parametrize("dim, field, shape", [(3, 2, (20, 18, 16)), (2, 1, (24, 20, 1)), (2, 0, (22, 26, 1))],
                         ids=["leveque_3d", "vortex_2d", "rotation_2d"])
def test_face_velocities_bitwise(cells, dim, field, shape):
    nx, ny, nz = shape
    o = O.Oracle(O.make_params(nx, dim=dim, ny=ny, nz=nz if dim == 3 else None))
    t, t2s = 0.37, 3.0
    o.velocity_faces(field, t, t2s)
    got = face_velocity(cells, o, field, t, t2s)
    for name, g in zip(("UF", "VF", "WF"), got):
        want = interior(o, name)
        assert np.array_equal(g, want), (name, np.max(np.abs(g - want)))


def run_wlic(cells, o, stencil=0):
    """One vof_wlic on the host cells from the oracle's current state
    (stencil: the cell update that loads its stencil up front)."""
    p = o.p
    o.surface_norms()   # what vof_wlic computes first, from the same phi
    vof = interior(o, "VOF").copy()
    args = [interior(o, n) for n in ("VOF_OLD", "UF", "VF", "WF", "NX", "NY")]
    n2 = interior(o, "NZ") if p.dim == 3 else None
    cells.hc_vof_wlic(p.dim, p.nx, p.ny, p.nz, p.dt, p.dx, p.dy, p.dz, ptr(vof), *(ptr(a) for a in args),
                      ptr(n2) if n2 is not None else None, stencil)
    return vof


@pytest.mark.parametrize("dim, shape", [(3, (24, 22, 20)), (2, (40, 36, 1))], ids=["sphere_3d", "circle_2d"])
def test_wlic_passes_bitwise(cells, dim, shape):
    nx, ny, nz = shape
    o = O.Oracle(O.make_params(nx, dim=dim, ny=ny, nz=nz if dim == 3 else None))
    field = 2 if dim == 3 else 1
    o.initial_phi(2 if dim == 3 else 1, 1)
    o.store_initial()
    for step in range(3):
        if step:
            o.clsvof_step(field, 3.0, 1)   # a transported state: more cut cells, both slopes
        o.velocity_faces(field, (step + 1) * o.p.dt, 3.0)
        o.field("VOF_OLD")[:] = o.field("VOF")
        got = run_wlic(cells, o)
        got_stencil = run_wlic(cells, o, stencil=1)
        o.vof_wlic()
        want = interior(o, "VOF")
        cut = np.count_nonzero((want > 1e-10) & (want < 1 - 1e-10))
        assert cut > 50
        assert np.array_equal(got, want), (step, np.max(np.abs(got - want)))
        assert np.array_equal(got_stencil, want), (step, np.max(np.abs(got_stencil - want)))


def bptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_ubyte))


def pass_inputs(cells, o):
    """vof as each WLIC pass of one vof_wlic reads it (before the x, y, z passes)."""
    p = o.p
    o.surface_norms()
    args = [interior(o, n) for n in ("VOF_OLD", "UF", "VF", "WF", "NX", "NY")]
    n2 = interior(o, "NZ") if p.dim == 3 else None
    out = []
    for npass in range(p.dim):
        v = interior(o, "VOF").copy()
        cells.hc_vof_wlic_n(p.dim, p.nx, p.ny, p.nz, p.dt, p.dx, p.dy, p.dz, ptr(v), *(ptr(a) for a in args),
                            ptr(n2) if n2 is not None else None, 0, npass)
        out.append(v)
    return out


def cut_line_flags(cells, o):
    p = o.p
    v = interior(o, "VOF")
    fx = np.zeros((p.nz, p.ny), np.uint8)
    fy = np.zeros((p.nz, p.nx), np.uint8)
    fz = np.zeros((p.ny, p.nx), np.uint8)
    cells.hc_cut_line_flags(ptr(v), p.nx, p.ny, p.nz, bptr(fx), bptr(fy), bptr(fz))
    return fx, fy, fz


def box_case(n):
    """An axis-aligned box of VOF 1 with faces on cell faces (no cut cell at
    the start: the first pass makes them) and its signed distance."""
    o = O.Oracle(O.make_params(n, dim=3))
    c = (np.arange(n + 6) - 2.5) / n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    lo, hi = 0.3, 0.6
    d = np.maximum.reduce([lo - x, x - hi, lo - y, y - hi, lo - z, z - hi])
    o.field("PHI")[:] = -d
    o.field("VOF")[:] = (d < 0).astype(float)
    return o


@pytest.mark.parametrize("case", ["sphere", "sphere_moved", "box", "circle_2d"])
def test_cut_lines_cover_every_cut_cell(cells, case):
    """The lines that k_cut_line_flags lists (wlic_may_cut on the vof at the
    step's start) hold every cell that any of the step's passes finds cut, so
    the normals solved on them are all the fluxes read."""
    if case == "box":
        o = box_case(24)
    else:
        dim = 2 if case == "circle_2d" else 3
        o = O.Oracle(O.make_params(40 if dim == 2 else 24, dim=dim))
        o.initial_phi(2 if dim == 3 else 1, 1)
        o.store_initial()
        if case == "sphere_moved":
            for _ in range(3):
                o.clsvof_step(2, 3.0, 1)
    p = o.p
    field = 2 if p.dim == 3 else 1
    for t in (0.1, 1.2):   # two velocity fields: the flow reverses in time
        o.velocity_faces(field, t, 3.0)
        o.field("VOF_OLD")[:] = o.field("VOF")
        fx, fy, fz = cut_line_flags(cells, o)
        ncut = 0
        for v in pass_inputs(cells, o):
            cut = (v >= 1e-10) & (v <= 1 - 1e-10)   # not (v > 1 - eps or v < eps)
            kk, jj, ii = np.nonzero(cut)
            ncut += kk.size
            assert np.all(fx[kk, jj] == 1) and np.all(fy[kk, ii] == 1)
            if p.dim == 3:
                assert np.all(fz[jj, ii] == 1)
        assert ncut > 0
        listed = fx.sum() + fy.sum() + (fz.sum() if p.dim == 3 else 0)
        total = p.ny * p.nz + p.nx * p.nz + (p.nx * p.ny if p.dim == 3 else 0)
        assert listed < total   # fewer line solves than every line

