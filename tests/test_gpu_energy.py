"""lsmp_energy_momentum (energy_and_momentum, output.f90:250-397) through the
C ABI against the oracle on the same fields: the eleven integrals to 1e-12
(FP64; the momenta, which can cancel, to 1e-12 of the velocity scale) and
1e-5 (FP32 bulk).  Cases: 3D and 2D, non-cubic, more than 4,096 lines per
direction (the one-lane CCD solve), lines of 600 cells (64 lanes per line),
the inverse density and viscosity branches, the VOF as the marker.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import load_lsmp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    dict(nx=40, ny=20, nz=36, dim=3),
    dict(nx=72, ny=66, nz=68, dim=3),
    dict(nx=600, ny=10, nz=8, dim=3),
    dict(nx=70, ny=50, dim=2),
    dict(nx=48, dim=3, rho_12=0.7, inverse=1, mu=0.3),
    dict(nx=48, dim=3, vof=True, mu=2.0),
    dict(nx=64, ny=48, dim=2, vof=True),
]


def ids(c):
    return "x".join(str(c.get(k)) for k in ("nx", "ny", "nz") if c.get(k)) + f"_{c['dim']}d" + \
        ("_inv" if c.get("inverse") else "") + ("_vof" if c.get("vof") else "")


def run(case, precision):
    lsmp = load_lsmp()
    c = dict(case)
    nx, dim = c.pop("nx"), c.pop("dim")
    vof, mu = c.pop("vof", False), c.pop("mu", 1.0)
    ny = c.pop("ny", nx)
    nz = c.pop("nz", nx) if dim == 3 else 1
    o = O.Oracle(O.make_params(nx, dim=dim, ny=ny, nz=nz, **c))
    s = lsmp.Solver(lsmp.make_params(nx, dim=dim, ny=ny, nz=nz, precision=precision, **c))
    o.initial_phi(2 if dim == 3 else 1, 1 if vof else 0)
    o.velocity(2 if dim == 3 else 1, 0.4, 3.0)
    s.upload_phi(o.field("PHI"))
    s.set_velocity(o.field("U"), o.field("V"), o.field("W"))
    if vof:
        s.upload_vof(o.field("VOF"))
    e_o = o.energy_momentum(mu, vof)
    e_g = s.energy_momentum(mu, vof)
    s.close()
    vscale = max(np.abs(o.interior(f)).max() for f in ("U", "V", "W"))
    return e_o, e_g, vscale


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_energy_momentum_fp64(case):
    say("energy fp64", case)
    e_o, e_g, vscale = run(case, 0)
    for k in e_o:
        tol = 1e-12 * (vscale if k in ("px", "py", "pz") else abs(e_o[k]))
        say(f"  {k}: oracle {e_o[k]:.15e} gpu {e_g[k]:.15e}")
        assert abs(e_g[k] - e_o[k]) <= tol + 1e-300, k


@pytest.mark.parametrize("case", [CASES[0], CASES[3], CASES[5]], ids=ids)
def test_energy_momentum_fp32(case):
    say("energy fp32", case)
    e_o, e_g, vscale = run(case, 1)
    for k in e_o:
        tol = 1e-5 * (vscale if k in ("px", "py", "pz") else abs(e_o[k]))
        assert abs(e_g[k] - e_o[k]) <= tol + 1e-300, (k, e_g[k], e_o[k])


def test_energy_needs_velocity():
    lsmp = load_lsmp()
    s = lsmp.Solver(lsmp.make_params(24, dim=3))
    o = O.Oracle(O.make_params(24, dim=3))
    o.initial_phi(2, 0)
    s.upload_phi(o.field("PHI"))
    with pytest.raises(RuntimeError):
        s.energy_momentum()
    s.close()
