"""The WLIC options against the default path on the GPU (not collected by
pytest until they have passed on an MI355X: run the file by naming it).

* LSMP_WLIC_STENCIL=1, the cell update that loads its stencil first: the
  same operations on the same values, so three merged CLSVOF steps agree to
  1e-15 (bit for bit unless the compiler contracts the two differently);
* LSMP_CUT_LINES=1, the normals solved only on the lines through cells a
  step can find cut: the same line solves on those lines (32 lanes per line
  instead of one lane where there are few), so the steps agree to 1e-12.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import load_lsmp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(n, dim, stencil, steps=3, var="LSMP_WLIC_STENCIL"):
    lsmp = load_lsmp()
    o = O.Oracle(O.make_params(n, dim=dim))
    o.initial_phi(2 if dim == 3 else 1, 1)
    old = os.environ.get(var)
    os.environ[var] = "1" if stencil else "0"
    try:
        s = lsmp.Solver(lsmp.make_params(n, dim=dim))
        s.upload_phi(o.field("PHI"))
        s.upload_vof(o.field("VOF"))   # the switches are read here
    finally:
        if old is None:
            del os.environ[var]
        else:
            os.environ[var] = old
    t = 0.0
    field = lsmp.VEL_LEVEQUE3D if dim == 3 else lsmp.VEL_VORTEX2D
    for it in range(1, steps + 1):
        t += s.params.dt
        s.clsvof_step(t, 3.0, field, iteration=it)
    out = s.download_vof(), s.download_phi()
    s.close()
    return out


@pytest.mark.parametrize("dim, n", [(3, 40), (2, 64)], ids=["3d", "2d"])
def test_stencil_update_matches_default(dim, n):
    say("wlic stencil", dim, n)
    va, pa = run(n, dim, False)
    vb, pb = run(n, dim, True)
    ev, ep = np.max(np.abs(va - vb)), np.max(np.abs(pa - pb))
    say(f"  max |vof diff| {ev:.3e}, max |phi diff| {ep:.3e}, bitwise {np.array_equal(va, vb)}")
    assert ev <= 1e-15 and ep <= 1e-15


@pytest.mark.parametrize("dim, n", [(3, 40), (2, 64)], ids=["3d", "2d"])
def test_cut_line_normals_match_default(dim, n):
    say("wlic cut lines", dim, n)
    va, pa = run(n, dim, False, var="LSMP_CUT_LINES")
    vb, pb = run(n, dim, True, var="LSMP_CUT_LINES")
    ev, ep = np.max(np.abs(va - vb)), np.max(np.abs(pa - pb))
    say(f"  max |vof diff| {ev:.3e}, max |phi diff| {ep:.3e}")
    assert ev <= 1e-12 and ep <= 1e-12

