"""Mass correction on band lines (lsmp_set_band_lines): the CCD normals are
solved only on the lines that cross |phi| < ls_wid, where the correction
delta |grad phi| and its sum are nonzero; the other lines add only their
mass and volume.  Bar: the drop-in tolerance (1e-12 relative) against solving
every line, step after step, in both precisions, with the listed lines solved
32 lanes per line (lines up to 540 cells) or one lane per line (longer), and
for an interface that crosses the domain's boundary lines."""
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


CASES = [
    dict(nx=72, ny=66, nz=68, dim=3, precision=0),
    dict(nx=72, ny=66, nz=68, dim=3, precision=1),
    dict(nx=72, ny=66, nz=68, dim=3, precision=0, corner=True),
    dict(nx=600, ny=560, dim=2, precision=0),
    dict(nx=600, ny=560, dim=2, precision=1),
    dict(nx=200, ny=180, dim=2, precision=0, corner=True),
]


def ids(c):
    dims = "x".join(str(c[k]) for k in ("nx", "ny", "nz") if k in c)
    return f"{dims}_{c['dim']}d_{'fp32' if c['precision'] else 'fp64'}" + ("_corner" if c.get("corner") else "")


def start_field(nx, ny, nz, dim, corner):
    """The oracle's initial sphere (disk), or one centred on a domain corner."""
    o = O.Oracle(O.make_params(nx, dim=dim, ny=ny, nz=nz))
    o.initial_phi(2 if dim == 3 else 1, 0)
    box = o.field("PHI").copy()
    if corner:
        nzb, nyb, nxb = box.shape
        x = ((np.clip(np.arange(nxb) - 3, 0, nx - 1) + 0.5) / nx)[None, None, :]
        y = ((np.clip(np.arange(nyb) - 3, 0, ny - 1) + 0.5) / ny)[None, :, None]
        r2 = (x - 0.04) ** 2 + (y - 0.03) ** 2
        if dim == 3:
            z = ((np.clip(np.arange(nzb) - 3, 0, nz - 1) + 0.5) / nz)[:, None, None]
            r2 = r2 + (z - 0.05) ** 2
        box = np.broadcast_to(0.3 - np.sqrt(r2), box.shape).copy()
    return box


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_band_lines_match_all_lines(case):
    lsmp = load_lsmp()
    dim = case["dim"]
    nx, ny = case["nx"], case["ny"]
    nz = case.get("nz", 1)
    box = start_field(nx, ny, nz, dim, case.get("corner", False))
    prm = lsmp.make_params(nx, dim=dim, ny=ny, nz=nz if dim == 3 else None, precision=case["precision"])
    field = lsmp.VEL_LEVEQUE3D if dim == 3 else lsmp.VEL_VORTEX2D
    t2s = 3.0 if dim == 3 else 8.0
    # FP32 bulk: both runs round phi to FP32 after every kernel, so they agree
    # to FP32 rounding of the field, not to 1e-12
    tol = TOL if case["precision"] == 0 else 1e-5
    mtol = 1e-13 if case["precision"] == 0 else 1e-8
    say("band lines", case)
    with lsmp.Solver(prm) as a, lsmp.Solver(prm) as b:
        a.set_band_lines(2)   # band lines always (the default takes them in 3D only)
        b.set_band_lines(0)
        for s in (a, b):
            s.upload_phi(box)
            s.store_phi0()
        imass, _ = a.mass_vol()
        # a correction that needs several iterations: shift phi by a tenth of a cell
        for s in (a, b):
            s.upload_phi(box + 0.1 * prm.dx)
        # the lines listed: those with a cell |phi| < ls_wid (1 + 1e-6), per direction
        ph = a.download_phi()
        ph = ph[3:-3, 3:-3, 3:-3] if dim == 3 else ph[3:4, 3:-3, 3:-3]
        band = np.abs(ph) < prm.ls_wid * (1.0 + 1.0e-6)   # (z, y, x)
        want = [int(band.any(axis=2).sum()), int(band.any(axis=1).sum()),
                int(band.any(axis=0).sum()) if dim == 3 else 0]
        ma, na = a.mass_fix(imass)
        mb, nb = b.mass_fix(imass)
        say(f"  listed lines {a.band_line_counts()} (expected {want}), {b.band_line_counts()} with band lines off")
        assert a.band_line_counts() == want
        assert b.band_line_counts() == [-1, -1, -1]
        e = rel(a.download_phi(), b.download_phi())
        say(f"  mass fix: iterations {na}/{nb}, mass {ma!r}/{mb!r}, rel diff {e:.3e}")
        assert abs(na - nb) <= 1 and abs(ma - mb) <= mtol * abs(mb) and e <= tol
        t = 0.0
        for step in range(3):
            t += prm.dt
            ra = a.mpls_step(t, t2s, field, imass)
            rb = b.mpls_step(t, t2s, field, imass)
            e = rel(a.download_phi(), b.download_phi())
            say(f"  step {step + 1}: pseudo/mass {ra} {rb}, rel diff {e:.3e}")
            assert ra[0] == rb[0] and abs(ra[1] - rb[1]) <= 1
            assert e <= tol
        ma, _ = a.mass_vol()
        assert abs(ma - imass) <= (1e-13 if case["precision"] == 0 else 1e-7) * imass


def test_band_lines_argument_checked():
    lsmp = load_lsmp()
    with lsmp.Solver(lsmp.make_params(8, dim=3)) as s:
        assert s.lib.lsmp_set_band_lines(s.h, 3) == -1   # LSMP_E_INVALID
        assert s.lib.lsmp_set_band_lines(s.h, -1) == -1
        assert s.lib.lsmp_set_band_lines(None, 1) == -1
        for mode in (0, 2, 1):
            s.set_band_lines(mode)


@pytest.mark.parametrize("case, listed", [(dict(nx=72, ny=66, nz=68, dim=3), True),
                                          (dict(nx=40, ny=36, nz=32, dim=3), False),
                                          (dict(nx=200, ny=180, dim=2), False)],
                         ids=["72x66x68_3d", "40x36x32_3d", "200x180_2d"])
def test_band_lines_default_where_they_pay(case, listed):
    """The default lists lines on 3D grids with more than 4096 lines in a
    direction, and solves every line elsewhere."""
    lsmp = load_lsmp()
    dim = case["dim"]
    box = start_field(case["nx"], case["ny"], case.get("nz", 1), dim, False)
    prm = lsmp.make_params(case["nx"], dim=dim, ny=case["ny"], nz=case.get("nz") if dim == 3 else None)
    with lsmp.Solver(prm) as s:
        s.upload_phi(box)
        imass, _ = s.mass_vol()
        s.upload_phi(box + 0.1 * prm.dx)
        s.mass_fix(imass)
        counts = s.band_line_counts()
        say("default band lines", case, counts)
        assert (min(counts[:dim]) > 0) if listed else counts == [-1, -1, -1]
