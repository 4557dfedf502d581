"""GPU checks of the multi-sphere case: the runner at a small size (every
sphere its own interface at the start, total mass conserved, per-interface
mass changes reported), and the per-interface partition of
lsmp_component_masses against scipy labels."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import PKG_DIR, load_lsmp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_multisphere_runner_small():
    load_lsmp()
    spec = importlib.util.spec_from_file_location("multisphere", os.path.join(PKG_DIR, "multisphere.py"))
    M = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(M)
    for form in ("fp32-learned", "fp64-plain"):
        # the phi > 0 cells of two spheres closer than about a cell (1/96) can
        # share a face: draw them 0.03 apart
        r = M.run(n=96, form=form, frac=0.01, count=12, gap=0.03, log=lambda m: print(m, flush=True))
        print({k: v for k, v in r.items() if not k.startswith("interface_mass")}, flush=True)
        assert r["steps"] > 0 and r["components_start"] == 12
        assert r["dM_over_M"] < (1e-6 if form.startswith("fp32") else 1e-12)
        assert r["spheres_compared"] >= 10 and r["interface_dM_over_M_max"] < 0.05
        assert abs(r["mass_in_components_end"] - 1) < 1e-3


def _grow(lab, band):
    """Layer-by-layer growth of labels (0 = none) into the band cells, each
    taking the smallest label among its labelled face neighbours."""
    big = np.iinfo(np.int64).max
    lab = lab.astype(np.int64)
    while True:
        L = np.pad(np.where(lab > 0, lab, big), 1, constant_values=big)
        m = np.minimum.reduce([L[:-2, 1:-1, 1:-1], L[2:, 1:-1, 1:-1], L[1:-1, :-2, 1:-1], L[1:-1, 2:, 1:-1],
                               L[1:-1, 1:-1, :-2], L[1:-1, 1:-1, 2:]])
        new = band & (lab == 0) & (m < big)
        if not new.any():
            return lab
        lab[new] = m[new]


def test_component_masses_core_and_band_against_scipy():
    """Two spheres 2 cells apart: their bands (phi > -ls_wid) touch, their
    cores (phi > 0) do not.  With phi_core = 0 the GPU keeps them apart and
    splits the shared band like scipy labels of the cores grown layer by
    layer in numpy; the masses match a numpy ls_mv sum over each part and add
    up to the solver's total.  With phi_core = phi_min they are one component."""
    ndimage = pytest.importorskip("scipy.ndimage")
    lsmp = load_lsmp()
    n = 40
    h = 1.0 / n
    x = (np.arange(-3, n + 3) + 0.5) / n
    z, y, xx = np.meshgrid(x, x, x, indexing="ij", sparse=True)
    spheres = [((0.3, 0.5, 0.5), 0.12), ((0.59, 0.5, 0.5), 0.12), ((0.3, 0.78, 0.78), 0.1)]
    box = np.full((n + 6,) * 3, -1.0)
    for (cx, cy, cz), r in spheres:
        box = np.maximum(box, r - np.sqrt((xx - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2))
    with lsmp.Solver(lsmp.make_params(n, dim=3)) as s:
        s.upload_phi(np.ascontiguousarray(box))
        total, _ = s.mass_vol()
        p = s.params
        phi = s.download_phi()[3:-3, 3:-3, 3:-3]
        core, ncore = ndimage.label(phi > 0)
        lab = _grow(core, phi > -p.ls_wid)
        xw = phi / p.ls_wid
        H = np.where(xw > 1 - 1e-12, 1.0, np.where(xw < -1 + 1e-12, 0.0,
                                                 0.5 * ((1 + xw) + np.sin(np.pi * xw) / np.pi)))
        H = np.clip(H, 0.0, 1.0)
        rho = p.rho_12 * H + (1 - H) if p.inverse else H + p.rho_12 * (1 - H)
        want = np.array([np.sum((rho * H)[lab == k]) * h ** 3 for k in range(1, ncore + 1)])
        cells = np.flatnonzero(lab.ravel() > 0)
        masses, comp = s.component_masses(-p.ls_wid, cells, phi_core=0.0)
        say(f"core components {ncore}; gpu {masses}; numpy {want}")
        assert ncore == 3 and masses.size == 3
        # same partition, numbered alike (both by the smallest cell index)
        assert np.array_equal(comp, lab.ravel()[cells] - 1)
        assert np.allclose(masses, want, rtol=1e-12, atol=0)
        assert abs(masses.sum() - total) <= 1e-12 * total
        merged, _ = s.component_masses(-p.ls_wid, cells)
        assert merged.size == 2
