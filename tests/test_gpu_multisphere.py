"""GPU smoke of the multi-sphere case runner at a small size: every sphere is
its own interface at the start, the total mass is conserved by the mass
correction, per-interface mass changes are reported for every sphere."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import PKG_DIR, load_lsmp

pytestmark = pytest.mark.gpu


def test_multisphere_runner_small():
    load_lsmp()
    spec = importlib.util.spec_from_file_location("multisphere", os.path.join(PKG_DIR, "multisphere.py"))
    M = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(M)
    for form in ("fp32-learned", "fp64-plain"):
        # at 96^3 the regions phi > -ls_wid of two spheres share a face when their
        # surfaces are closer than 3.5 dx = 0.036: draw them 0.045 apart
        r = M.run(n=96, form=form, frac=0.01, count=12, gap=0.045, log=lambda m: print(m, flush=True))
        print(r, flush=True)
        assert r["steps"] > 0 and r["components_start"] == 12
        assert r["dM_over_M"] < (1e-6 if form.startswith("fp32") else 1e-12)
        assert np.isfinite(r["interface_dM_over_M_max"]) and r["interface_dM_over_M_max"] < 0.2
