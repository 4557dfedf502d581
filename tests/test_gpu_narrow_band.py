"""Narrow band (lsmp_set_narrow_band, the chosen forms' option): advection
clamps phi to [-W, W], solves only the lines crossing the tube |phi| < W and
updates only the cells whose three lines cross it; reinitialisation updates
only the workgroups of 32 x 8 x 8 cells that hold a tube cell.  Checked:
every other cell keeps its (clamped) value bit for bit, and near the
interface (|phi| < 3 dx) the field follows the full-domain method to the
tolerances written in each test."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import load_lsmp

pytestmark = pytest.mark.gpu

W = 10.0          # tube half-width in cells
TILE = (32, 8, 8)   # reinitialisation workgroup of the narrow band: x, y, z cells


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [dict(nx=72, ny=66, nz=68, precision=0), dict(nx=72, ny=66, nz=68, precision=1)]


def ids(c):
    return f"{c['nx']}x{c['ny']}x{c['nz']}_{'fp32' if c['precision'] else 'fp64'}"


def setup(case, nb, scale=1.0):
    lsmp = load_lsmp()
    o = O.Oracle(O.make_params(case["nx"], dim=3, ny=case["ny"], nz=case["nz"]))
    o.initial_phi(2, 0)
    box = o.field("PHI").copy() * scale
    prm = lsmp.make_params(case["nx"], dim=3, ny=case["ny"], nz=case["nz"], precision=case["precision"])
    s = lsmp.Solver(prm)
    s.set_narrow_band(W if nb else 0.0)
    s.upload_phi(box)
    return s, prm


def interior(b):
    return b[3:-3, 3:-3, 3:-3]


def tube_cells(phi, w):
    """Cells whose x, y and z lines all hold a cell with |phi| < w or a face
    neighbour of one (z, y, x)."""
    core = np.abs(phi) < w
    band = core.copy()
    for ax in range(3):
        for sh in (1, -1):
            r = np.roll(core, sh, axis=ax)
            edge = [slice(None)] * 3
            edge[ax] = 0 if sh == 1 else -1
            r[tuple(edge)] = False   # no wrap-around
            band |= r
    fx = band.any(axis=2)[:, :, None]
    fy = band.any(axis=1)[:, None, :]
    fz = band.any(axis=0)[None, :, :]
    return fx & fy & fz


def tile_cells(phi, w):
    """Cells of the workgroups that hold a cell with |phi| < w (z, y, x)."""
    band = np.abs(phi) < w
    nz, ny, nx = band.shape
    tx, ty, tz = TILE
    pad = np.zeros((-(-nz // tz) * tz, -(-ny // ty) * ty, -(-nx // tx) * tx), bool)
    pad[:nz, :ny, :nx] = band
    t = pad.reshape(pad.shape[0] // tz, tz, pad.shape[1] // ty, ty, pad.shape[2] // tx, tx).any(axis=(1, 3, 5))
    full = np.repeat(np.repeat(np.repeat(t, tz, axis=0), ty, axis=1), tx, axis=2)
    return full[:nz, :ny, :nx]


def width(prm):
    """The tube's half-width: W times the largest cell size."""
    return W * max(prm.dx, prm.dy, prm.dz)


# at 192^3 more than 4096 lines cross the tube in each direction: the tiled
# line solve, with the tube mask in its z pass
@pytest.mark.parametrize("case", CASES + [dict(nx=192, ny=192, nz=192, precision=0)], ids=ids)
def test_advect_keeps_cells_outside_the_tube(case):
    lsmp = load_lsmp()
    a, prm = setup(case, True)
    b, _ = setup(case, False)
    say("narrow band advect", case)
    t = 0.3
    before = interior(a.download_phi())
    for s in (a, b):
        s.set_velocity_analytic(lsmp.VEL_LEVEQUE3D, t, 3.0)
        s.rk3_advect()
    pa, pb = interior(a.download_phi()), interior(b.download_phi())
    w = width(prm)
    tube = tube_cells(before, w)
    # outside the tube: the field clamped to +-W, in the solver's precision
    wt = float(np.float32(w)) if case["precision"] else w
    want = np.where(np.abs(before) > w, np.sign(before) * wt, before)
    say(f"  tube cells {tube.mean():.3f} of the box; changed outside: {np.count_nonzero(pa[~tube] != want[~tube])}")
    assert np.array_equal(pa[~tube], want[~tube])
    assert np.count_nonzero(pa[tube] != before[tube]) > 0.5 * tube.sum()
    band = np.abs(before) < 3 * prm.dx
    d = np.max(np.abs(pa[band] - pb[band])) / prm.dx
    say(f"  band |narrow - full| / dx = {d:.3e}")
    # the clamped field beyond W changes the right-hand side along each solved
    # line, and the implicit line solves carry a fraction of that (about 0.55
    # per cell) over the W - 3 cells to the band: 1.4e-4 dx measured
    assert d <= 1e-3
    a.close()
    b.close()


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_reinit_keeps_workgroups_outside_the_tube(case):
    # a field 1.3 times a distance: the reinitialisation moves every tube cell
    a, prm = setup(case, True, scale=1.3)
    b, _ = setup(case, False, scale=1.3)
    say("narrow band reinit", case)
    before = interior(a.download_phi())
    _, na = a.reinit(1)
    _, nb = b.reinit(1)
    pa, pb = interior(a.download_phi()), interior(b.download_phi())
    tiles = tile_cells(before, width(prm))
    say(f"  pseudo-steps {na}/{nb}; tile cells {tiles.mean():.3f} of the box")
    assert np.array_equal(pa[~tiles], before[~tiles])
    band = np.abs(before) < 3 * prm.dx
    d = np.max(np.abs(pa[band] - pb[band])) / prm.dx
    say(f"  band |narrow - full| / dx = {d:.3e}")
    # 30 pseudo-steps move information 3 dx, outward from the interface: the
    # frozen workgroups, at least W - 3 cells from the band, reach it only
    # through the small weights of the WENO-HJ stencils' downwind cells
    # (7e-6 dx measured)
    assert d <= 1e-4
    a.close()
    b.close()


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_mpls_steps_near_the_full_domain(case):
    lsmp = load_lsmp()
    a, prm = setup(case, True)
    b, _ = setup(case, False)
    say("narrow band MPLS steps", case)
    for s in (a, b):
        s.store_phi0()
    imass, _ = a.mass_vol()
    t = 0.0
    for step in range(4):
        t += prm.dt
        ra = a.mpls_step(t, 3.0, lsmp.VEL_LEVEQUE3D, imass)
        rb = b.mpls_step(t, 3.0, lsmp.VEL_LEVEQUE3D, imass)
        pa, pb = interior(a.download_phi()), interior(b.download_phi())
        band = np.abs(pb) < 3 * prm.dx
        d = np.max(np.abs(pa[band] - pb[band])) / prm.dx
        say(f"  step {step + 1}: pseudo/mass {ra} {rb}, band |narrow - full| / dx = {d:.3e}")
        assert d <= 1e-3   # as for one advection, per step (1.4e-4 dx measured)
    ma, _ = a.mass_vol()
    assert abs(ma - imass) <= (1e-13 if case["precision"] == 0 else 1e-7) * imass
    assert abs(a.interface_l1() - b.interface_l1()) <= 1e-3 * b.interface_l1()
    a.close()
    b.close()


def test_narrow_band_arguments():
    lsmp = load_lsmp()
    with lsmp.Solver(lsmp.make_params(16, dim=2)) as s:
        assert s.lib.lsmp_set_narrow_band(s.h, 10.0) == -1   # 3D only
        s.set_narrow_band(0.0)
    with lsmp.Solver(lsmp.make_params(16, dim=3)) as s:
        assert s.lib.lsmp_set_narrow_band(s.h, 2.0) == -1    # narrower than the stencils
        assert s.lib.lsmp_set_narrow_band(s.h, -1.0) == -1
        assert s.lib.lsmp_set_narrow_band(s.h, float("nan")) == -1
        s.set_narrow_band(8.0)
        s.set_narrow_band(0.0)
