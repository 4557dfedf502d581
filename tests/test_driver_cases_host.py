"""The Fortran driver's initial conditions (fortran/lsmp_cases.f90: phi and
the 30^3 sub-sampled VOF of problem_init.f90 for Disk, Vortex2D and Vortex,
with the zero-gradient ghosts) equal the oracle's bit for bit, built with
the driver's own compiler and flags.  Skipped without a Fortran compiler."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
CELLS = os.path.join(HERE, "host_cells")
CHECK = os.path.join(CELLS, "_build", "cases_check")


@pytest.fixture(scope="module")
def checker():
    if shutil.which("flang") is None:
        pytest.skip("no Fortran compiler in this image")
    subprocess.run(["make", "-s", "-C", CELLS, "cases"], check=True)
    return CHECK


@pytest.mark.parametrize("case, shape, n", [("disk", 0, 32), ("vortex2d", 1, 40), ("vortex", 2, 24)])
def test_initial_phi_and_vof_match_oracle(checker, tmp_path, case, shape, n):
    out = tmp_path / f"{case}.bin"
    subprocess.run([checker, case, str(n), str(out)], check=True, timeout=120)
    dim = 3 if case == "vortex" else 2
    o = O.Oracle(O.make_params(n, dim=dim))
    o.initial_phi(shape, 1)
    want_phi, want_vof = o.field("PHI"), o.field("VOF")
    got = np.fromfile(out, dtype=np.float64)
    assert got.size == 2 * want_phi.size
    phi = got[:want_phi.size].reshape(want_phi.shape)
    vof = got[want_phi.size:].reshape(want_phi.shape)
    assert np.count_nonzero((want_vof > 0) & (want_vof < 1)) > 50
    assert np.array_equal(phi, want_phi), np.max(np.abs(phi - want_phi))
    assert np.array_equal(vof, want_vof), np.max(np.abs(vof - want_vof))
