"""The WLIC pass's stencil-first cell update (LSMP_WLIC_STENCIL=1) against the
default one on the GPU (not collected by pytest until it has passed on an
MI355X: run it by naming the file).  Both perform the same operations on the
same values, so three merged CLSVOF steps must agree to 1e-15 (bit for bit
unless the compiler contracts the two differently)."""
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


def run(n, dim, stencil, steps=3):
    lsmp = load_lsmp()
    o = O.Oracle(O.make_params(n, dim=dim))
    o.initial_phi(2 if dim == 3 else 1, 1)
    old = os.environ.get("LSMP_WLIC_STENCIL")
    os.environ["LSMP_WLIC_STENCIL"] = "1" if stencil else "0"
    try:
        s = lsmp.Solver(lsmp.make_params(n, dim=dim))
        s.upload_phi(o.field("PHI"))
        s.upload_vof(o.field("VOF"))   # the stencil switch is read here
    finally:
        if old is None:
            del os.environ["LSMP_WLIC_STENCIL"]
        else:
            os.environ["LSMP_WLIC_STENCIL"] = old
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
