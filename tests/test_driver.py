"""The Fortran driver (ISO_C_BINDING over the C ABI) on decks in the
reference's format: it reproduces the reference's main loop and output files,
and its result equals the same step sequence issued from Python."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from conftest import PKG_DIR, ROOT, load_lsmp

DRIVER = os.path.join(PKG_DIR, "lib", "lsmp_driver")
DECKS = os.path.join(ROOT, "tests", "decks")


def test_driver_built():
    if shutil.which("flang") is None and shutil.which("amdflang") is None:
        pytest.skip("no Fortran compiler in this image")
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"


def test_driver_rejects_unknown_case(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    r = subprocess.run([DRIVER, "clsvof"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "unknown case" in r.stdout


def run_driver(tmp_path, case, deck, steps):
    r = subprocess.run([DRIVER, case, os.path.join(DECKS, deck), str(steps)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"finished: steps\s+(\d+)\s+L1\s+(\S+)\s+max \|dM\|/M\s+(\S+)", r.stdout)
    assert m, r.stdout
    return int(m.group(1)), float(m.group(2)), float(m.group(3))


def python_path(n, dim, field, t2s, steps, rdt_cfl=0.1, shape=2, init_btn0=False):
    """The same calls from Python: what the driver must reproduce."""
    lsmp = load_lsmp()
    o = O.Oracle(O.make_params(n, dim=dim, rdt_cfl=rdt_cfl))
    o.initial_phi(shape, 1 if shape == 0 else 0)
    s = lsmp.Solver(lsmp.make_params(n, dim=dim, rdt_cfl=rdt_cfl))
    s.upload_phi(o.field("PHI"))
    if init_btn0:
        s.reinit(0)
    s.store_phi0()
    imass, _ = s.mass_vol()
    t = 0.0
    dmax = 0.0
    for it in range(1, steps + 1):
        t = t + s.params.dt
        s.mpls_step(t, t2s, field, imass, iteration=it)
        m, _ = s.mass_vol()
        dmax = max(dmax, abs(imass - m) / imass)
    l1 = s.interface_l1()
    s.close()
    return l1, dmax


@pytest.mark.gpu
def test_driver_vortex3d_matches_python_path(tmp_path):
    steps, l1, dmax = run_driver(tmp_path, "vortex", "vortex_mpls_16.txt", 4)
    assert steps == 4
    l1p, dmaxp = python_path(16, 3, 2, 3.0, 4)
    assert abs(l1 - l1p) <= 1e-12 * l1p
    assert dmax < 1e-12
    rows = np.loadtxt(tmp_path / "out" / "Vortex16_MassLoss.plt", skiprows=1)
    assert rows.shape == (4, 5)
    assert np.all(np.abs(rows[:, 1]) < 1e-9)      # percent mass loss
    text = (tmp_path / "Vortex_Results.txt").read_text()
    assert "Method: MPLS" in text and "LS Int:" in text


@pytest.mark.gpu
def test_driver_vortex2d_matches_python_path(tmp_path):
    steps, l1, dmax = run_driver(tmp_path, "vortex2d", "vortex2d_mpls_32.txt", 5)
    l1p, _ = python_path(32, 2, 1, 8.0, 5, shape=1)
    assert abs(l1 - l1p) <= 1e-12 * l1p


@pytest.mark.gpu
def test_driver_disk_with_btn0_init(tmp_path):
    """Disk: phi0 = 2 VOF - 1 from 30^3 sub-samples, then redis(0)
    (Disk/src/setup/problem_init.f90:45-74, 103)."""
    steps, l1, dmax = run_driver(tmp_path, "disk", "disk_mpls_32.txt", 3)
    l1p, _ = python_path(32, 2, 0, 1.0, 3, rdt_cfl=0.5, shape=0, init_btn0=True)
    assert abs(l1 - l1p) <= 1e-12 * max(l1p, 1e-30)
    assert dmax < 1e-12


@pytest.mark.gpu
def test_driver_classical_ls_loses_mass(tmp_path):
    """Method 1 (classical_ls, interface_solver.f90:15-23) has no mass fix."""
    steps, l1, dmax = run_driver(tmp_path, "vortex", "vortex_ls_16.txt", 4)
    assert dmax > 1e-8


@pytest.mark.gpu
def test_driver_energy_rows(tmp_path):
    """The driver's `energy` option: Energy, Momentum and Position rows every
    step (output.f90:33-36) from lsmp_energy_momentum with the deck's scales
    (tree.f90:289-318, selector 1), against the same calls from Python; the
    rows are printed with four significant digits."""
    steps = 3
    r = subprocess.run([DRIVER, "vortex", os.path.join(DECKS, "vortex_mpls_16.txt"), str(steps), "fp64", "crweno",
                        "energy"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    en = np.loadtxt(tmp_path / "out" / "Vortex16_Energy.plt", skiprows=1, ndmin=2)
    mom = np.loadtxt(tmp_path / "out" / "Vortex16_Momentum.plt", skiprows=1, ndmin=2)
    pos = np.loadtxt(tmp_path / "out" / "Vortex16_Position.plt", skiprows=1, ndmin=2)
    assert en.shape == (steps, 7) and mom.shape == (steps, 4) and pos.shape == (steps, 4)
    # deck vortex_mpls_16.txt: rho_1, mu_1, sigma, g, Re, Fr, We, mu_12, btn_sf
    rho_1, mu_1, g, re, fr, we, mu_12 = 995.0, 0.0007972, 9.81, 67.27, 1.0, 16.0, 0.01
    L = ((re * mu_1 / rho_1) ** 2 * g / fr) ** (1.0 / 3.0)
    unit = rho_1 * g * L ** 4
    lsmp = load_lsmp()
    o = O.Oracle(O.make_params(16, dim=3))
    o.initial_phi(2, 0)
    s = lsmp.Solver(lsmp.make_params(16, dim=3))
    s.upload_phi(o.field("PHI"))
    s.store_phi0()
    imass, _ = s.mass_vol()
    t, ev = 0.0, -1.0
    for it in range(1, steps + 1):
        t = t + s.params.dt
        s.mpls_step(t, 3.0, 2, imass, iteration=it)
        e = s.energy_momentum(mu_12)
        es, ek, ep = e["A"] * fr / we * unit, e["Ke"] * fr * unit, e["Po"] * unit
        if ev < 0:
            e0, ev = ek + ep + es, 0.0
        else:
            ev = ev + e["Q"] * s.params.dt * fr / re * unit
        row = en[it - 1]
        assert row[1:5] == pytest.approx([ek, ep, es, ev], rel=1e-3, abs=1e-300)
        assert row[6] == pytest.approx(e0, rel=1e-3)
        assert mom[it - 1, 1:] == pytest.approx([e["px"], e["py"], e["pz"]], rel=1e-3, abs=1e-8)
        assert pos[it - 1, 1:] == pytest.approx([e["xc"], e["yc"], e["zc"]], rel=1e-3)
    s.close()
