"""CPU checks of the 512^3 multi-sphere case set-up (seeded spheres and the
initial level-set box): no GPU, small grids, brute force."""
import importlib.util
import os

import numpy as np

from conftest import PKG_DIR


def _ms():
    spec = importlib.util.spec_from_file_location("multisphere", os.path.join(PKG_DIR, "multisphere.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_spheres_seeded_disjoint_and_inside():
    M = _ms()
    c, r = M.seeded_spheres(64, seed=2026)
    assert c.shape == (64, 3) and r.shape == (64,)
    assert r.min() >= 0.02 and r.max() <= 0.06
    assert np.all(c - r[:, None] >= 0.01 - 1e-15) and np.all(c + r[:, None] <= 0.99 + 1e-15)
    d = np.linalg.norm(c[:, None] - c[None], axis=-1) - r[:, None] - r[None]
    np.fill_diagonal(d, 1.0)
    assert d.min() > 0.01
    c2, r2 = M.seeded_spheres(64, seed=2026)
    assert np.array_equal(c, c2) and np.array_equal(r, r2)


def test_spheres_box_matches_brute_force():
    M = _ms()
    c, r = M.seeded_spheres(6, rmin=0.08, rmax=0.15, seed=7)
    n = 24
    box = M.spheres_box(n, c, r)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    want = np.full((n, n, n), -0.25)
    for cc, rr in zip(c, r):
        want = np.maximum(want, rr - np.sqrt((X - cc[0]) ** 2 + (Y - cc[1]) ** 2 + (Z - cc[2]) ** 2))
    assert np.allclose(box[3:-3, 3:-3, 3:-3], want, atol=1e-15, rtol=0)
    # ghosts are copies of the nearest interior cell
    assert np.array_equal(box[0, 5, 7], box[3, 5, 7]) and np.array_equal(box[5, -1, 7], box[5, -4, 7])
    # each sphere's centre cell is inside (phi > 0)
    ijk = np.clip(np.floor(c * n).astype(int), 0, n - 1)
    assert np.all(box[ijk[:, 2] + 3, ijk[:, 1] + 3, ijk[:, 0] + 3] > 0)


def test_tracked_centres_return_after_one_period():
    """The LeVeque field reverses at T/2, so its flow map over [0, T] is the
    identity: tracked centres move far and come back, with the first-order
    error of a velocity frozen over each step (halving with dt)."""
    M = _ms()
    c, _ = M.seeded_spheres(64, seed=2026)
    errs = []
    for n in (96, 192):
        dt, T = 0.1 / n, 3.0
        p, t, far = c.copy(), 0.0, 0.0
        nsteps = int(round(T / dt))
        for it in range(nsteps):
            p = M.track_points(p, t, dt, T)
            t += dt
            if it == nsteps // 2:
                far = np.max(np.linalg.norm(p - c, axis=1))
        assert far > 0.5
        errs.append(np.max(np.linalg.norm(p - c, axis=1)))
    assert errs[0] < 3e-3 and 1.8 < errs[0] / errs[1] < 2.2, errs


def test_velocity_is_divergence_free():
    M = _ms()
    rng = np.random.default_rng(3)
    p = rng.uniform(0.05, 0.95, size=(50, 3))
    h = 1e-5
    div = np.zeros(50)
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        div += (M.leveque_velocity(p + e, 0.4, 3.0)[:, a] - M.leveque_velocity(p - e, 0.4, 3.0)[:, a]) / (2 * h)
    assert np.max(np.abs(div)) < 1e-8
