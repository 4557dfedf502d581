"""Host logic of the Fortran driver, without a GPU: lsmp_driver built against a
stand-in for liblsmp that records the entry points it is called through
(tests/host_cells/mock_lsmp.c) runs the reference's main-loop order for each
method (main.f90:17-51; interface_solver.f90:15-23, 123-129, 234-244) and
writes the reference's result blocks (output.f90:169-231).  Skipped without
a Fortran compiler."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CELLS = os.path.join(HERE, "host_cells")
DRIVER = os.path.join(CELLS, "_build", "mock", "lsmp_driver")
DECKS = os.path.join(HERE, "decks")

INIT = ["lsmp_init", "lsmp_set_advection_scheme", "lsmp_upload_phi"]


@pytest.fixture(scope="module")
def driver():
    if shutil.which("flang") is None:
        pytest.skip("no Fortran compiler in this image")
    subprocess.run(["make", "-s", "-C", CELLS, "cases"], check=True)
    return DRIVER


def run(driver, tmp_path, case, deck, steps, extra=()):
    log = tmp_path / "calls.log"
    env = dict(os.environ, LSMP_MOCK_LOG=str(log))
    r = subprocess.run([driver, case, os.path.join(DECKS, deck), str(steps)] + list(extra), cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return log.read_text().split(), r.stdout


def test_mpls_order(driver, tmp_path):
    calls, _ = run(driver, tmp_path, "vortex", "vortex_mpls_16.txt", 3)
    step = ["lsmp_set_velocity_analytic", "lsmp_rk3_advect", "lsmp_reinit(1)", "lsmp_mass_fix", "lsmp_mass_vol"]
    want = INIT + ["lsmp_store_phi0", "lsmp_mass_vol", "lsmp_synchronize"] + step * 3 \
        + ["lsmp_synchronize", "lsmp_interface_l1", "lsmp_finalize"]
    assert calls == want
    assert "Method: MPLS" in (tmp_path / "Vortex_Results.txt").read_text()


def test_energy_option_adds_the_diagnostic_each_step(driver, tmp_path):
    """With `energy` the driver calls energy_and_momentum after the step's
    mass sums (output.f90:33) and writes one row per step to the Energy,
    Momentum and Position files."""
    calls, _ = run(driver, tmp_path, "vortex", "vortex_mpls_16.txt", 3, ("fp64", "crweno", "energy"))
    step = ["lsmp_set_velocity_analytic", "lsmp_rk3_advect", "lsmp_reinit(1)", "lsmp_mass_fix", "lsmp_mass_vol",
            "lsmp_energy_momentum"]
    want = INIT + ["lsmp_store_phi0", "lsmp_mass_vol", "lsmp_synchronize"] + step * 3 \
        + ["lsmp_synchronize", "lsmp_interface_l1", "lsmp_finalize"]
    assert calls == want
    for f, cols in (("Energy", 7), ("Momentum", 4), ("Position", 4)):
        lines = (tmp_path / "out" / f"Vortex16_{f}.plt").read_text().splitlines()
        assert len(lines) == 4 and all(len(l.split()) == cols for l in lines[1:])


def test_classical_ls_has_no_mass_fix(driver, tmp_path):
    calls, _ = run(driver, tmp_path, "vortex", "vortex_ls_16.txt", 2)
    assert "lsmp_mass_fix" not in calls and calls.count("lsmp_rk3_advect") == 2
    assert "Method: LS" in (tmp_path / "Vortex_Results.txt").read_text()


def test_disk_starts_with_btn0_reinitialisation(driver, tmp_path):
    calls, _ = run(driver, tmp_path, "disk", "disk_mpls_32.txt", 1)
    assert calls[:5] == INIT + ["lsmp_reinit(0)", "lsmp_store_phi0"]


@pytest.mark.parametrize("case, deck", [("vortex", "vortex_clsvof_16.txt"), ("vortex2d", "vortex2d_clsvof_32.txt")])
def test_clsvof_order_and_outputs(driver, tmp_path, case, deck):
    calls, out = run(driver, tmp_path, case, deck, 3)
    step = ["lsmp_set_velocity_analytic", "lsmp_set_face_velocity_analytic", "lsmp_store_vof_old", "lsmp_vof_wlic",
            "lsmp_rk3_advect", "lsmp_clsvof_recon", "lsmp_mass_vol", "lsmp_vof_mass_vol"]
    want = INIT + ["lsmp_store_phi0", "lsmp_mass_vol", "lsmp_upload_vof", "lsmp_store_vof0", "lsmp_vof_mass_vol",
                   "lsmp_synchronize"] + step * 3 \
        + ["lsmp_synchronize", "lsmp_interface_l1", "lsmp_vof_interface_l1", "lsmp_finalize"]
    assert calls == want
    text = (tmp_path / "Vortex_Results.txt").read_text()
    assert "Method: CLSVOF" in text and "VOF Mass:" in text and "VOF Int:" in text and "LS Mass:" not in text
    assert "VOF: L1" in out
    name = [f for f in os.listdir(tmp_path / "out") if f.endswith("_MassLoss.plt")][0]
    rows = np.loadtxt(tmp_path / "out" / name, skiprows=1)
    assert rows.shape == (3, 5)


@pytest.mark.parametrize("config, method", [("vortex64_clsvof", 3), ("vortex128", 2), ("disk200", 2)])
def test_config_decks_run_through_the_driver(driver, tmp_path, monkeypatch, config, method):
    """configs.py's decks parse in the driver; the reference's own Vortex deck
    (CLSVOF, t2red = -1) runs without a reconstruction."""
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mass-preserving-level-set-method_amd"))
    import configs
    log = tmp_path / "calls.log"
    monkeypatch.setenv("LSMP_MOCK_LOG", str(log))
    res = configs.run(config, "fp64", max_steps=3, driver=driver)
    calls = log.read_text().split()
    assert res["steps"] == 3 and res["method"] == method
    if method == 3:
        assert "vof_l1" in res
        assert calls.count("lsmp_vof_wlic") == 3 and "lsmp_clsvof_recon" not in calls
    else:
        assert calls.count("lsmp_mass_fix") == 3
