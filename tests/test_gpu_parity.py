"""GPU parity: the HIP library, called through its C ABI, against the CPU
oracle on identical seeded inputs.

FP64 bar: max |phi_gpu - phi_oracle| <= 1e-12 * max |phi_oracle| per call and
per step (the drop-in tolerance).  FP32 bulk form: compared with the FP64
path through the L1 shape error and |dM|/M.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import load_lsmp

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(nx, ny=None, nz=None, dim=3, precision=0, **kw):
    """Oracle and GPU solver on the same grid and parameters."""
    lsmp = load_lsmp()
    ny = nx if ny is None else ny
    nz = (nx if nz is None else nz) if dim == 3 else 1
    op = O.make_params(nx, dim=dim, ny=ny, nz=nz, **kw)
    gp = lsmp.make_params(nx, dim=dim, ny=ny, nz=nz, precision=precision, **kw)
    for f in ("dx", "dy", "dz", "dt", "rdt", "ls_wid"):
        assert getattr(op, f) == getattr(gp, f)
    return O.Oracle(op), lsmp.Solver(gp)


def rel_err(gpu_box, orc_box):
    return np.max(np.abs(gpu_box - orc_box)) / np.max(np.abs(orc_box))


def load_state(o, s, shape=2, perturb=False, field=2, t=0.3, t2s=3.0):
    """Initial field and velocity from the oracle, copied to the GPU."""
    o.initial_phi(shape, 0)
    if perturb:
        phi = o.field("PHI")
        nzb, nyb, nxb = phi.shape
        x = ((np.arange(nxb) - 2.5) / o.p.nx)[None, None, :]
        y = ((np.arange(nyb) - 2.5) / o.p.ny)[None, :, None]
        phi *= 0.6 + 0.8 * x + 0.3 * np.sin(3 * y)
        o.bc_field("PHI")
    o.velocity(field, t, t2s)
    s.upload_phi(o.field("PHI"))
    s.set_velocity(o.field("U"), o.field("V"), o.field("W"))


CASES = [
    dict(nx=24, dim=3),
    dict(nx=40, ny=20, nz=36, dim=3),
    dict(nx=48, dim=2),
    dict(nx=70, ny=50, dim=2),
    # odd line lengths: the two halves of a twisted line solve differ by a row
    dict(nx=23, ny=17, nz=19, dim=3),
    dict(nx=45, ny=31, dim=2),
    # x lines of 300 cells: the partitioned solves with 16 (CRWENO) and 17 (CCD) rows per lane
    dict(nx=300, ny=36, dim=2),
    # x lines of 640 cells: the CCD and UCCD coefficient tables exceed 64 KB and
    # are read from global memory instead of LDS
    dict(nx=640, ny=40, dim=2),
]


def ids(c):
    return "x".join(str(c.get(k)) for k in ("nx", "ny", "nz") if c.get(k)) + f"_{c['dim']}d"


# more than 4096 lines per direction: the one-lane-per-line CCD kernel, with
# the mass correction's line lists and flags, in 3D
CASES_MANY_LINES = [dict(nx=72, ny=66, nz=68, dim=3)]

# advection passes of 4096 lines or more take the tiled line solve: lines of
# 64..72 cells in 3D, x lines of 300 cells (16 rows per lane), and 2D y lines
# of 40 cells whose last workgroup of 8 lines holds 4
CASES_TILED = CASES_MANY_LINES + [dict(nx=300, ny=64, nz=64, dim=3), dict(nx=4100, ny=40, dim=2)]

# lines of 514..1025 cells take the long-line form of the tiled solve (16 rows
# per lane, 4 lines per workgroup) whatever their count, and lines of 541..1084
# cells the CCD solve with 64 lanes per line: x lines of 1025 cells
# (the longest), 3D x lines of 600 cells (the x lines of 640 cells in CASES as
# well); advection alone on strided y lines of 520 cells (dt follows dx, so y
# may not be much finer than x)
CASES_LONG = [dict(nx=1025, ny=12, dim=2), dict(nx=600, ny=10, nz=8, dim=3)]
CASES_LONG_ADVECT = [dict(nx=700, ny=520, dim=2)]


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_upload_download_roundtrip(case):
    o, s = pair(**case)
    say("roundtrip", case)
    o.initial_phi(2 if case["dim"] == 3 else 1, 0)
    s.upload_phi(o.field("PHI"))
    assert np.array_equal(s.download_phi(), o.field("PHI"))
    s.close()


@pytest.mark.parametrize("case", CASES + CASES_TILED + CASES_LONG + CASES_LONG_ADVECT, ids=ids)
def test_advect_parity(case):
    o, s = pair(**case)
    say("advect", case)
    load_state(o, s, shape=2 if case["dim"] == 3 else 1, field=2 if case["dim"] == 3 else 1)
    for _ in range(2):
        o.rk3_advect()
        s.rk3_advect()
        e = rel_err(s.download_phi(), o.field("PHI"))
        say("  rel err", e)
        assert e <= TOL
    s.close()


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_advect_uccd_parity(case):
    """The UCCD advection variant (level_set_rk3_source_ccd) against the
    oracle's UCCD RHS, two RK3 steps, 2D and 3D, non-cubic grids."""
    o, s = pair(**case)
    lsmp = load_lsmp()
    o.set("scheme", 1)
    s.set_advection_scheme(lsmp.SCHEME_UCCD)
    say("advect uccd", case)
    load_state(o, s, shape=2 if case["dim"] == 3 else 1, field=2 if case["dim"] == 3 else 1, perturb=True)
    for _ in range(2):
        o.rk3_advect()
        s.rk3_advect()
        e = rel_err(s.download_phi(), o.field("PHI"))
        say("  rel err", e)
        assert e <= TOL
    # switching back restores the default scheme on the same handle
    s.set_advection_scheme(lsmp.SCHEME_CRWENO)
    o.set("scheme", 0)
    o.rk3_advect()
    s.rk3_advect()
    assert rel_err(s.download_phi(), o.field("PHI")) <= TOL
    s.close()


def test_uccd_fp32_close_to_fp64():
    lsmp = load_lsmp()
    n = 48
    outs = []
    for prec in (lsmp.FP64, lsmp.FP32):
        o = O.Oracle(O.make_params(n, dim=3))
        o.initial_phi(2, 0)
        o.velocity(2, 0.3, 3.0)
        s = lsmp.Solver(lsmp.make_params(n, dim=3, precision=prec))
        s.set_advection_scheme(lsmp.SCHEME_UCCD)
        s.upload_phi(o.field("PHI"))
        s.set_velocity(o.field("U"), o.field("V"), o.field("W"))
        for _ in range(5):
            s.rk3_advect()
        outs.append(s.download_phi())
        s.close()
    d = np.max(np.abs(outs[0] - outs[1]))
    say("uccd fp32 vs fp64 max diff", d)
    assert np.isfinite(outs[1]).all() and d < 1e-5


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_analytic_velocity_matches_uploaded(case):
    """Device-evaluated velocity (velocity.f90) gives the same step as the
    oracle's host velocity uploaded as boxes."""
    o, s = pair(**case)
    say("velocity", case)
    field = 2 if case["dim"] == 3 else 1
    load_state(o, s, shape=2 if case["dim"] == 3 else 1, field=field, t=0.45, t2s=3.0)
    s.set_velocity_analytic(field, 0.45, 3.0)
    o.rk3_advect()
    s.rk3_advect()
    assert rel_err(s.download_phi(), o.field("PHI")) <= TOL
    s.close()


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_reinit_parity(case):
    o, s = pair(**case)
    say("reinit", case)
    load_state(o, s, shape=2 if case["dim"] == 3 else 1, perturb=True)
    no = o.redis(1, -1.0)
    err_g, ng = s.reinit(1)
    say("  pseudo-steps", no, ng, "red_error", o.get("red_error"), err_g)
    assert ng == no
    assert abs(err_g - o.get("red_error")) <= 1e-9 * o.get("red_error")
    e = rel_err(s.download_phi(), o.field("PHI"))
    say("  rel err", e)
    assert e <= TOL
    s.close()


@pytest.mark.parametrize("case", [CASES[0], CASES[2]], ids=ids)
def test_reinit_btn0_parity(case):
    """Disk-style initialisation: raw-phi sign, phi /= max|grad|, no lambda
    (level_set_rk3_redis.f90:111, 116-174, 378-396)."""
    o, s = pair(rdt_cfl=0.5, **case)
    say("reinit btn0", case)
    o.initial_phi(0 if case["dim"] == 2 else 2, 1)
    s.upload_phi(o.field("PHI"))
    T = 12 * o.p.rdt
    no = o.redis(0, T)
    _, ng = s.reinit(0, T)
    assert ng == no
    e = rel_err(s.download_phi(), o.field("PHI"))
    say("  rel err", e)
    assert e <= TOL
    s.close()


@pytest.mark.parametrize("case", CASES + CASES_MANY_LINES + CASES_LONG, ids=ids)
def test_mass_fix_parity(case):
    o, s = pair(**case)
    say("mass fix", case)
    o.initial_phi(2 if case["dim"] == 3 else 1, 0)
    o.store_initial()
    imass = o.get("imass")
    o.field("PHI")[:] += 0.3 * o.p.dx
    s.upload_phi(o.field("PHI"))
    m0, _ = s.mass_vol()
    o.ls_mv()
    assert abs(m0 - o.get("mass")) <= 1e-13 * o.get("mass")
    no = o.mass_fix()
    mg, ng = s.mass_fix(imass)
    say("  iterations", no, ng)
    assert abs(ng - no) <= 1
    e = rel_err(s.download_phi(), o.field("PHI"))
    say("  rel err", e)
    assert e <= TOL
    mass, vol = s.mass_vol()
    assert abs(mass - imass) / imass < 1e-13
    s.close()


@pytest.mark.parametrize("case", CASES + CASES_MANY_LINES + CASES_LONG, ids=ids)
def test_mpls_steps_parity(case):
    """Full merged-MPLS steps (interface_solver.f90:123-129) with reinit every
    step, velocity at t^{n+1} (main.f90:23-29)."""
    o, s = pair(**case)
    say("mpls steps", case)
    dim = case["dim"]
    field = 2 if dim == 3 else 1
    o.initial_phi(2 if dim == 3 else 1, 0)
    o.store_initial()
    s.upload_phi(o.field("PHI"))
    s.store_phi0()
    imass = o.get("imass")
    m, _ = s.mass_vol()
    assert abs(m - imass) <= 1e-13 * imass
    t = 0.0
    for step in range(3):
        t = t + o.p.dt
        o.mpls_step(field, 3.0, 1)
        npseudo, nmass = s.mpls_step(t, 3.0, field, imass)
        e = rel_err(s.download_phi(), o.field("PHI"))
        say(f"  step {step + 1}: rel err {e:.3e}  pseudo {npseudo}/{o.get('redis_iters'):.0f}"
            f"  mass iters {nmass}/{o.get('mass_iters'):.0f}")
        assert e <= TOL
    l1g = s.interface_l1()
    l1o = o.interface_l1()
    assert abs(l1g - l1o) <= 1e-9 * max(l1o, 1e-30)
    s.close()


def test_minimum_size_grids():
    for case in (dict(nx=4, dim=3), dict(nx=4, ny=5, dim=2), dict(nx=5, ny=4, nz=6, dim=3)):
        say("minimum", case)
        o, s = pair(**case)
        load_state(o, s, shape=2 if case["dim"] == 3 else 1, perturb=True,
                   field=2 if case["dim"] == 3 else 1)
        o.rk3_advect(); s.rk3_advect()
        assert rel_err(s.download_phi(), o.field("PHI")) <= TOL
        o.redis(1, -1.0); s.reinit(1)
        assert rel_err(s.download_phi(), o.field("PHI")) <= TOL
        s.close()


def test_mass_fix_minimum_size_grids():
    """The CCD normals of the mass correction on the smallest grids, where
    most of the 32 segments of a line are empty."""
    for case in (dict(nx=4, dim=3), dict(nx=4, ny=5, dim=2), dict(nx=5, ny=4, nz=6, dim=3), dict(nx=7, dim=2)):
        say("minimum mass fix", case)
        o, s = pair(**case)
        o.initial_phi(2 if case["dim"] == 3 else 1, 0)
        o.store_initial()
        imass = o.get("imass")
        o.field("PHI")[:] += 0.1 * o.p.dx
        s.upload_phi(o.field("PHI"))
        no = o.mass_fix()
        _, ng = s.mass_fix(imass)
        e = rel_err(s.download_phi(), o.field("PHI"))
        say(f"  iterations {no}/{ng}  rel err {e:.3e}")
        assert abs(ng - no) <= 1 and e <= TOL
        s.close()


def test_empty_band_is_an_error():
    """No cell inside the band: the reference divides by zero
    (mass_preserving_level_set.f90:46); the library reports it."""
    lsmp = load_lsmp()
    o, s = pair(nx=16, dim=3)
    o.field("PHI")[:] = 1.0
    s.upload_phi(o.field("PHI"))
    with pytest.raises(lsmp.LsmpError, match="-4"):
        s.mass_fix(0.5)
    s.close()


def test_call_order_errors():
    lsmp = load_lsmp()
    s = lsmp.Solver(lsmp.make_params(16, dim=3))
    with pytest.raises(lsmp.LsmpError, match="-5"):
        s.rk3_advect()
    with pytest.raises(lsmp.LsmpError, match="-5"):
        s.interface_l1()
    s.close()


def test_fp32_bulk_form_close_to_fp64():
    """FP32 advection / reinit / CCD with FP64 mass sums: after several steps
    the shape differs from the FP64 path by far less than the discretisation
    error and mass holds to 1e-6."""
    lsmp = load_lsmp()
    n = 32
    say("fp32 vs fp64", n)
    o = O.Oracle(O.make_params(n, dim=3))
    o.initial_phi(2, 0)
    o.store_initial()
    imass = o.get("imass")
    res = {}
    for prec in (0, 1):
        s = lsmp.Solver(lsmp.make_params(n, dim=3, precision=prec))
        s.upload_phi(o.field("PHI"))
        s.store_phi0()
        t = 0.0
        for _ in range(8):
            t += s.params.dt
            s.mpls_step(t, 3.0, 2, imass)
        mass, _ = s.mass_vol()
        res[prec] = (s.download_phi(), s.interface_l1(), abs(mass - imass) / imass)
        s.close()
    (p64, l64, dm64), (p32, l32, dm32) = res[0], res[1]
    say("  L1 fp64", l64, "fp32", l32, "dM/M", dm64, dm32)
    assert dm64 < 1e-12
    assert dm32 < 1e-6
    assert np.max(np.abs(p32 - p64)) < 1e-3
    assert abs(l32 - l64) <= 0.1 * l64 + 1e-5


def test_full_size_step_properties():
    """At the headline size: one full step conserves mass to 1e-12 and is
    bitwise reproducible (fixed-order reductions)."""
    lsmp = load_lsmp()
    n = 256
    say("256^3 step properties")
    o = O.Oracle(O.make_params(32, dim=3))
    outs = []
    for rep in range(2):
        s = lsmp.Solver(lsmp.make_params(n, dim=3))
        box = np.empty(s.shape)
        zb = (np.arange(-3, n + 3) + 0.5) / n
        z, y, x = np.meshgrid(zb, zb, zb, indexing="ij", sparse=True)
        c, r = float(np.float32(0.35)), float(np.float32(0.15))
        box[:] = -np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) + r
        s.upload_phi(box)
        imass, _ = s.mass_vol()
        npseudo, nmass = s.mpls_step(s.params.dt, 3.0, 2, imass)
        mass, _ = s.mass_vol()
        say(f"  rep {rep}: pseudo {npseudo} mass iters {nmass} dM/M {abs(mass - imass) / imass:.2e}")
        assert abs(mass - imass) / imass < 1e-12
        outs.append(s.download_phi())
        s.close()
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0], outs[1])
