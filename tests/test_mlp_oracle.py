"""Pins of the learned-reinitialisation oracle (CPU).

The network has no reference counterpart, so its oracle is pinned by
mathematics and by library routines: central finite differences for the
backward pass, torch's bfloat16 conversion for the rounding, torch.optim.Adam
for the optimiser step, closed forms for the stencil features.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402


def random_params(rng, scale=1.0):
    P = np.zeros(M.NPARAM)
    P[M.OFF["W1"]:M.OFF["b1"]] = rng.normal(size=M.NF * M.NH) * np.sqrt(1 / 126) * scale
    P[M.OFF["b1"]:M.OFF["W2"]] = rng.normal(size=M.NH) * 0.1
    P[M.OFF["W2"]:M.OFF["b2"]] = rng.normal(size=M.NH * M.NH) * np.sqrt(1 / 64) * scale
    P[M.OFF["b2"]:M.OFF["w3"]] = rng.normal(size=M.NH) * 0.1
    P[M.OFF["w3"]:M.OFF["b3"]] = rng.normal(size=M.NH) * 0.3
    P[M.OFF["b3"]] = 0.05
    return P


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10 ** rng.uniform(-8, 8, 20000),
                        np.float32([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -3.0e38, 1e-40, np.inf])])
    # exact ties: 1 + 2^-8 (rounds down to even) and 1 + 3 * 2^-8 (rounds up)
    ties = np.float32([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8)])
    x = np.concatenate([x, ties]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = M.bf16(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert M.bf16(np.float32([1 + 2 ** -8]))[0] == 1.0
    assert M.bf16(np.float32([1 + 3 * 2 ** -8]))[0] == np.float32(1 + 2 ** -6)


@pytest.mark.parametrize("wc", [0.0, 0.2])
def test_backward_matches_finite_differences(wc):
    rng = np.random.default_rng(1)
    P = random_params(rng)
    X = rng.normal(size=(11, M.NF))
    Y = rng.normal(size=11) * 0.3
    L, g = M.loss_grad(P, X, Y, emulate=False, weight_c=wc)
    idx = np.concatenate([rng.choice(M.OFF["b1"], 60, replace=False),
                          np.arange(M.OFF["b1"], M.OFF["b1"] + 8),
                          M.OFF["W2"] + rng.choice(M.NH * M.NH, 40, replace=False),
                          np.arange(M.OFF["b2"], M.NPARAM)])
    h = 1e-6
    for p in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        fd = (M.loss_grad(Pp, X, Y, emulate=False, weight_c=wc)[0]
              - M.loss_grad(Pm, X, Y, emulate=False, weight_c=wc)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-7 + 1e-6 * abs(fd), (p, fd, g[p])


def test_backward_detects_planted_mistake():
    """A dropped SiLU derivative in the first layer fails the same check."""
    rng = np.random.default_rng(2)
    P = random_params(rng)
    X = rng.normal(size=(5, M.NF))
    Y = rng.normal(size=5)
    _, g = M.loss_grad(P, X, Y, emulate=False)
    orig = M.dsilu
    calls = []

    def wrong(z):
        calls.append(1)
        return orig(z) if len(calls) == 1 else np.ones_like(z)

    M.dsilu = wrong
    try:
        _, gw = M.loss_grad(P, X, Y, emulate=False)
    finally:
        M.dsilu = orig
    p = M.OFF["b1"] + 3
    h = 1e-6
    Pp, Pm = P.copy(), P.copy()
    Pp[p] += h
    Pm[p] -= h
    fd = (M.loss_grad(Pp, X, Y, emulate=False)[0] - M.loss_grad(Pm, X, Y, emulate=False)[0]) / (2 * h)
    assert abs(fd - g[p]) < 1e-6 * max(1, abs(fd))
    assert abs(fd - gw[p]) > 1e-4 * max(1e-3, abs(fd))


def test_emulated_pass_is_close_to_exact():
    rng = np.random.default_rng(3)
    P = random_params(rng)
    X = rng.normal(size=(64, M.NF))
    Y = rng.normal(size=64) * 0.1
    o_e, _ = M.forward(P, X, emulate=True)
    o_x, _ = M.forward(P, X, emulate=False)
    # split-BF16 forward: about 16 significant bits
    assert np.max(np.abs(o_e - o_x)) < 1e-3 * np.max(np.abs(o_x))
    L_e, g_e = M.loss_grad(P, X, Y, emulate=True)
    L_x, g_x = M.loss_grad(P, X, Y, emulate=False)
    assert abs(L_e - L_x) < 0.05 * L_x
    assert np.linalg.norm(g_e - g_x) < 0.05 * np.linalg.norm(g_x)


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    P0 = rng.normal(size=257).astype(np.float32)
    grads = [rng.normal(size=257).astype(np.float32) * s for s in (1.0, 0.01, 3.0)]
    t = torch.nn.Parameter(torch.from_numpy(P0.copy()))
    opt = torch.optim.Adam([t], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    P, m, v = P0.copy(), np.zeros_like(P0), np.zeros_like(P0)
    for step, g in enumerate(grads, start=1):
        t.grad = torch.from_numpy(g.copy())
        opt.step()
        P, m, v = M.adam(P, m, v, g, 1e-3, step)
        assert np.allclose(P, t.detach().numpy(), rtol=0, atol=2e-7)


def test_features_closed_form_on_linear_field():
    nz, ny, nx = 9, 8, 10
    h = 0.125
    a, b, c, d = 0.3, -0.7, 0.5, 0.11
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    phi = (a * i + b * j + c * k) * h + d
    cells = np.array([(4 * ny + 3) * nx + 5, 0])
    X = M.features(phi, cells, 1.0 / h)
    q = 0
    for dk in range(-2, 3):
        for dj in range(-2, 3):
            for di in range(-2, 3):
                assert X[0, q] == pytest.approx(a * di + b * dj + c * dk, abs=1e-6)
                # cell 0: offsets below 0 clamp to the boundary cell
                assert X[1, q] == pytest.approx(a * max(di, 0) + b * max(dj, 0) + c * max(dk, 0), abs=1e-6)
                q += 1
    assert X[0, 125] == pytest.approx(phi[4, 3, 5] / h, rel=1e-6)
    # central differences are exact on a linear field away from the boundary
    gm = np.sqrt(a * a + b * b + c * c)
    assert X[0, 126] == pytest.approx(gm - 1, abs=1e-6)
    assert X[0, 127] == pytest.approx(phi[4, 3, 5] / h * (1 / gm - 1), rel=1e-5)
    # at cell 0 the clamped stencil gives (8 f(1) - f(2) + ... ) / 12 per axis
    g0 = np.array([(-(2 * s) + 8 * s) / 12 for s in (a, b, c)])
    assert X[1, 126] == pytest.approx(np.linalg.norm(g0) - 1, abs=1e-6)


def test_band_is_ordered_and_complete():
    rng = np.random.default_rng(5)
    phi = rng.normal(size=(6, 7, 8))
    cells = M.band(phi, 0.5)
    assert np.all(np.diff(cells) > 0)
    assert set(cells.tolist()) == set(np.flatnonzero(np.abs(phi) < 0.5).tolist())


def test_minibatch_indices_in_range_and_spread():
    idx = M.minibatch(M.step_seed(3, 1), 1000, 4096)
    assert idx.min() >= 0 and idx.max() < 1000
    assert len(np.unique(idx)) > 900
