"""The Fortran driver's method 3 (CLSVOF) on the GPU.

* The driver on a method-3 deck (Vortex 16^3, Vortex2D 32^2) equals the same
  calls issued from Python (1e-12 on the LS and VOF shape errors), writes
  "Method: CLSVOF" and the VOF rows, and keeps the VOF volume.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import PKG_DIR, ROOT, load_lsmp

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG_DIR, "lib", "lsmp_driver")
DECKS = os.path.join(ROOT, "tests", "decks")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def start(n, dim):
    lsmp = load_lsmp()
    o = O.Oracle(O.make_params(n, dim=dim))
    o.initial_phi(2 if dim == 3 else 1, 1)
    s = lsmp.Solver(lsmp.make_params(n, dim=dim))
    s.upload_phi(o.field("PHI"))
    s.upload_vof(o.field("VOF"))
    return lsmp, o, s


def run_driver(tmp_path, case, deck, steps):
    say("driver", case, deck, steps)
    r = subprocess.run([DRIVER, case, os.path.join(DECKS, deck), str(steps)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    say(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"finished: steps\s+(\d+)\s+L1\s+(\S+)", r.stdout)
    mv = re.search(r"VOF: L1\s+(\S+)\s+max \|dM\|/M\s+(\S+)", r.stdout)
    assert m and mv, r.stdout
    return int(m.group(1)), float(m.group(2)), float(mv.group(1)), float(mv.group(2))


def python_path(n, dim, field, t2s, steps):
    lsmp, o, s = start(n, dim)
    s.store_phi0()
    s.store_vof0()
    t = 0.0
    for it in range(1, steps + 1):
        t += s.params.dt
        s.clsvof_step(t, t2s, field, iteration=it)
    out = s.interface_l1(), s.vof_interface_l1()
    s.close()
    return out


@pytest.mark.parametrize("case, deck, n, dim, field, t2s",
                         [("vortex", "vortex_clsvof_16.txt", 16, 3, 2, 3.0),
                          ("vortex2d", "vortex2d_clsvof_32.txt", 32, 2, 1, 8.0)], ids=["vortex", "vortex2d"])
def test_driver_clsvof_matches_python_path(tmp_path, case, deck, n, dim, field, t2s):
    steps, l1, l1v, dmv = run_driver(tmp_path, case, deck, 4)
    assert steps == 4
    l1p, l1vp = python_path(n, dim, field, t2s, 4)
    say(f"  L1 {l1!r} / {l1p!r}, VOF L1 {l1v!r} / {l1vp!r}, VOF max |dM|/M {dmv:.2e}")
    assert abs(l1 - l1p) <= 1e-12 * l1p
    assert abs(l1v - l1vp) <= 1e-12 * l1vp
    assert dmv < 1e-10
    text = (tmp_path / "Vortex_Results.txt").read_text()
    assert "Method: CLSVOF" in text and "VOF Int:" in text
