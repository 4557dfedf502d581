"""GPU checks of the learned reinitialisation against its CPU oracle
(oracle/mlp_oracle.py) and against the plain reinitialisation.

Tolerances: the kernels multiply BF16 values exactly and accumulate in FP32
in an order the oracle does not copy, so sums agree to FP32 rounding; in the
backward pass a BF16 rounding that lands on the other side of a tie then
moves one term by one BF16 ulp (2^-8 relative).  Gradients: norm-wise 2e-3
relative, per entry 1e-3 of the largest entry.  The forward pass is split
BF16 (about 16 significant bits): network outputs within 1e-5 h per cell.
"""
import os
import sys

import numpy as np
import pytest

from conftest import load_lsmp, ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def random_params(rng):
    P = np.zeros(M.NPARAM, dtype=np.float32)
    P[M.OFF["W1"]:M.OFF["b1"]] = rng.normal(size=M.NF * M.NH) * np.sqrt(1 / 126)
    P[M.OFF["b1"]:M.OFF["W2"]] = rng.normal(size=M.NH) * 0.1
    P[M.OFF["W2"]:M.OFF["b2"]] = rng.normal(size=M.NH * M.NH) * np.sqrt(1 / 64)
    P[M.OFF["b2"]:M.OFF["w3"]] = rng.normal(size=M.NH) * 0.1
    P[M.OFF["w3"]:M.OFF["b3"]] = rng.normal(size=M.NH) * 0.3
    P[M.OFF["b3"]] = 0.05
    return P


def sphere_box(n, c=0.5, r=0.3, dim=3, warp=0.0):
    x = (np.arange(-3, n + 3) + 0.5) / n
    if dim == 3:
        z, y, xx = np.meshgrid(x, x, x, indexing="ij", sparse=True)
        d = np.sqrt((xx - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
        box = r - d
        if warp:
            box = box * (1 + warp * np.sin(7 * xx) * np.cos(5 * y))
    else:
        y, xx = np.meshgrid(x, x, indexing="ij", sparse=True)
        box = np.broadcast_to(r - np.sqrt((xx - c) ** 2 + (y - c) ** 2), (7, n + 6, n + 6))
    return np.ascontiguousarray(box, dtype=np.float64)


def interior(box):
    return box[3:-3, 3:-3, 3:-3]


def test_param_layout_and_roundtrip():
    lsmp = load_lsmp()
    rng = np.random.default_rng(0)
    with lsmp.Net(seed=7) as net:
        assert net.nparam == M.NPARAM
        P0 = net.get_params()
        assert np.isfinite(P0).all() and np.abs(P0[:M.OFF["b1"]]).max() <= np.sqrt(3 / 126) + 1e-7
        P = random_params(rng)
        net.set_params(P)
        assert np.array_equal(net.get_params(), P)


@pytest.mark.parametrize("n", [64, 200, 1000])
def test_forward_backward_match_oracle(n):
    """Loss and all 12481 gradient entries of the MFMA kernels against the
    oracle's BF16-emulating backward pass (ragged batch sizes included)."""
    lsmp = load_lsmp()
    rng = np.random.default_rng(n)
    P = random_params(rng)
    X = (rng.normal(size=(n, M.NF)) * 1.5).astype(np.float32)
    X[:, 126:] = 0
    Y = (rng.normal(size=n) * 0.2).astype(np.float32)
    say(f"eval_grad n={n}")
    with lsmp.Net() as net:
        net.set_params(P)
        loss, g = net.eval_grad(X, Y)
    L, go = M.loss_grad(P.astype(np.float64), X, Y, emulate=True)
    say(f"  loss gpu {loss:.8e} oracle {L:.8e}")
    assert abs(loss - L) <= 1e-4 * L
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    worst = np.max(np.abs(g - go)) / np.max(np.abs(go))
    say(f"  grad rel norm err {err:.2e}, worst entry {worst:.2e}")
    assert err < 2e-3
    assert worst < 1e-3
    # each block of the gradient separately (a transposed block fails here)
    for a, b in (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "w3"), ("w3", "b3")):
        sl = slice(M.OFF[a], M.OFF[b])
        assert np.linalg.norm(g[sl] - go[sl]) <= 3e-3 * np.linalg.norm(go[sl]) + 1e-9, a


def test_train_steps_match_oracle_adam():
    """Two training steps (hashed minibatches, Adam with bias correction, the
    second with the relative weighting) against the oracle's minibatch draw,
    backward pass and Adam step."""
    lsmp = load_lsmp()
    rng = np.random.default_rng(11)
    nsamp, batch, lr, seed = 300, 128, 1e-3, 5
    P = random_params(rng)
    X = (rng.normal(size=(nsamp, M.NF)) * 1.5).astype(np.float32)
    X[:, 126:] = 0
    Y = (rng.normal(size=nsamp) * 0.2).astype(np.float32)
    say("train 2 steps")
    with lsmp.Net(capacity=nsamp) as net:
        net.set_params(P)
        net.add_samples(X, Y)
        assert net.dataset_size() == nsamp
        loss1 = net.train(1, batch=batch, lr=lr, seed=seed)
        P1 = net.get_params()
        loss2 = net.train(1, batch=batch, lr=lr, seed=seed, weight_c=0.1)
        P2 = net.get_params()
    m = np.zeros(M.NPARAM, np.float32)
    v = np.zeros(M.NPARAM, np.float32)
    Po = P.copy()
    for step, (loss_gpu, P_gpu, wc) in enumerate(((loss1, P1, 0.0), (loss2, P2, 0.1)), start=1):
        idx = M.minibatch(M.step_seed(seed, step), nsamp, batch)
        L, g = M.loss_grad(Po.astype(np.float64), X[idx], Y[idx], emulate=True, weight_c=wc)
        Pn, m, v = M.adam(Po, m, v, g, lr, step)
        say(f"  step {step}: loss gpu {loss_gpu:.6e} oracle {L:.6e}")
        assert abs(loss_gpu - L) <= 1e-4 * L
        # an Adam step moves each parameter by about lr * sign(g); entries
        # whose gradient is at rounding level may move either way
        big = np.abs(g) > 1e-2 * np.abs(g).max()
        assert np.max(np.abs((P_gpu - Po)[big] - (Pn - Po)[big])) <= 0.02 * lr
        assert np.max(np.abs(P_gpu - Po)) <= 1.01 * lr * np.sqrt(step) + 1e-7
        Po = P_gpu.astype(np.float32)


@pytest.mark.parametrize("precision", [0, 1])
def test_network_pass_matches_oracle(precision):
    """lsmp_net_apply (band list, stencil features, inference) against the
    oracle on a warped sphere, FP64 and FP32 fields."""
    lsmp = load_lsmp()
    n = 40
    rng = np.random.default_rng(3)
    P = random_params(rng)
    P[M.OFF["w3"]:] *= 0.1
    box = sphere_box(n, r=0.27, warp=0.15)
    say(f"net apply n={n} precision={precision}")
    with lsmp.Solver(lsmp.make_params(n, dim=3, precision=precision)) as s, lsmp.Net() as net:
        net.set_params(P)
        s.upload_phi(box)
        pre = interior(s.download_phi()).copy()
        nb = net.apply(s, 6.0)
        out = interior(s.download_phi())
    h = 1.0 / n
    pre_t = pre if precision == 0 else pre.astype(np.float32)
    want = M.apply(P.astype(np.float64), pre_t, h, 6.0, 1.5 * h)
    cells = M.band(pre_t, np.asarray(6.0 * h, dtype=pre_t.dtype))
    assert nb == cells.size
    d = np.abs(out - want)
    inband = np.zeros(pre.size, dtype=bool)
    inband[cells] = True
    inband = inband.reshape(pre.shape)
    say(f"  band {nb}, max |gpu - oracle| / h = {d.max() / h:.2e}")
    assert d.max() <= 1e-5 * h
    assert np.array_equal(out[~inband], pre[~inband])


def test_network_pass_without_band_count():
    """lsmp_net_apply with n_band = NULL only queues the pass (the band's
    length stays on the device): the same field bit for bit as the call that
    reads the length back, on fresh solvers (the pass's buffers are allocated
    on the first call)."""
    lsmp = load_lsmp()
    n = 32
    P = random_params(np.random.default_rng(5))
    P[M.OFF["w3"]:] *= 0.1
    box = sphere_box(n, r=0.25, warp=0.1)
    outs = []
    with lsmp.Net() as net:
        net.set_params(P)
        for read_count in (True, False):
            with lsmp.Solver(lsmp.make_params(n, dim=3, precision=1)) as s:
                s.upload_phi(box)
                pre = s.download_phi()
                if read_count:
                    nb = net.apply(s, 6.0)
                else:
                    assert net.lib.lsmp_net_apply(net.h, s.h, 6.0, None) == 0
                outs.append(s.download_phi())
    say(f"net apply without the count: band {nb}")
    assert nb > 0
    assert np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0], pre)   # the pass moved some band cells


def test_collected_pairs_match_plain_reinit():
    """lsmp_net_collect stores the features of the pre-reinit band and the
    plain method's correction: the loss of an lr = 0 step over them equals the
    oracle's loss on pairs rebuilt from phi before and after."""
    lsmp = load_lsmp()
    n = 32
    rng = np.random.default_rng(9)
    P = random_params(rng)
    P[M.OFF["w3"]:] *= 0.05
    box = sphere_box(n, r=0.3, warp=0.2)
    say("collect")
    with lsmp.Solver(lsmp.make_params(n, dim=3)) as s, lsmp.Net(capacity=200000) as net:
        net.set_params(P)
        s.upload_phi(box)
        pre = interior(s.download_phi()).copy()
        got = net.collect(s, 6.0)
        post = interior(s.download_phi()).copy()
        loss = net.train(1, batch=2048, lr=0.0, seed=3)
        assert np.array_equal(net.get_params(), P)
    h = 1.0 / n
    cells = M.band(pre, 6.0 * h)
    assert got == cells.size
    X = M.features(pre, cells, np.float32(1.0 / h), np.float32(1.0 / (12.0 * h)))
    Y = ((post.ravel()[cells] - pre.ravel()[cells]).astype(np.float32) * np.float32(1.0 / h))
    idx = M.minibatch(M.step_seed(3, 1), got, 2048)
    L, _ = M.loss_grad(P.astype(np.float64), X[idx], Y[idx], emulate=True)
    say(f"  pairs {got} loss gpu {loss:.6e} oracle {L:.6e}")
    assert abs(loss - L) <= 1e-3 * L


def test_residual_matches_oracle_gradient():
    """lsmp_reinit_residual against the mean | |grad phi| - 1 | of the
    oracle's Godunov gradient (level_set_rk3_redis.f90:266-369) on the band."""
    lsmp = load_lsmp()
    n = 32
    box = sphere_box(n, r=0.3, warp=0.2)
    o = O.Oracle(O.make_params(n, dim=3))
    o.field("PHI")[:] = box
    o.redis(1, -1.0)
    with lsmp.Solver(lsmp.make_params(n, dim=3)) as s:
        s.upload_phi(box)
        s.reinit(1)
        res = s.reinit_residual(3.0)
    o.redis_gradient()
    phi = o.interior("PHI")
    g = o.interior("GRAD")
    m = np.abs(phi) < 3.0 / n
    want = np.mean(np.abs(g[m] - 1.0))
    say(f"residual gpu {res:.12e} oracle {want:.12e}")
    assert abs(res - want) <= 1e-9 * want


def test_fallback_equals_plain_reinit():
    """With a residual threshold no field meets, the learned path must hand
    back exactly the plain method's result for the same input."""
    lsmp = load_lsmp()
    n = 32
    box = sphere_box(n, r=0.3, warp=0.2)
    with lsmp.Solver(lsmp.make_params(n, dim=3)) as a, lsmp.Solver(lsmp.make_params(n, dim=3)) as b, \
            lsmp.Net(seed=3) as net:
        a.upload_phi(box)
        b.upload_phi(box)
        _, n_plain = a.reinit(1)
        res, fb, n_l = b.reinit_learned(net, 6.0, 2, 1e-12, 3.0, lsmp.FALLBACK_RESTART)
        assert fb
        assert n_l == n_plain + 2
        assert np.array_equal(a.download_phi(), b.download_phi())
        res2, fb2, n2 = b.reinit_learned(net, 6.0, 2, 10.0, 3.0, lsmp.FALLBACK_RESTART)
        assert not fb2 and n2 == 2


def test_training_reduces_loss():
    """A short run of the project's own training kernels on pairs from the
    plain method lowers the loss well below its starting value."""
    lsmp = load_lsmp()
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "learned", os.path.join(ROOT, "mass-preserving-level-set-method_amd", "learned.py"))
    L = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(L)
    say("train from plain pairs")
    with lsmp.Net(capacity=400000, seed=2) as net:
        n = L.generate(net, sizes=(48,), shapes=3, chain=2, max_adv=4, seed=1, log=say)
        assert n > 1000
        l0 = net.train(1, batch=4096, lr=0.0, seed=9)
        hist = L.train(net, steps=600, batch=4096, lr=2e-3, lr_end=2e-4, chunk=200, seed=4, log=say)
    assert hist[-1][2] < 0.2 * l0


def test_continue_fallback_stops_within_plain_budget():
    """The continuation fallback runs plain pseudo-steps in blocks until the
    residual passes: with an unreachable threshold it spends exactly the
    plain method's pseudo-step count, with a loose one only the 2 sweeps."""
    lsmp = load_lsmp()
    n = 32
    box = sphere_box(n, r=0.3, warp=0.2)
    with lsmp.Solver(lsmp.make_params(n, dim=3)) as a, lsmp.Solver(lsmp.make_params(n, dim=3)) as b, \
            lsmp.Net(seed=3) as net:
        net.set_params(np.zeros(net.nparam, np.float32))
        a.upload_phi(box)
        _, n_plain = a.reinit(1)
        b.upload_phi(box)
        res, fb, n_c = b.reinit_learned(net, 6.0, 2, 1e-12, 1.5, lsmp.FALLBACK_CONTINUE)
        say(f"continue: plain {n_plain} pseudo-steps, continuation {n_c}, residual {res:.3e}")
        assert fb and n_c == n_plain
        # a zero network leaves phi unchanged, so the continuation with the full
        # budget is the plain method itself (same sign field, same steps)
        assert np.array_equal(a.download_phi(), b.download_phi())
        b.upload_phi(box)
        res, fb, n_c = b.reinit_learned(net, 6.0, 2, 10.0, 1.5, lsmp.FALLBACK_CONTINUE)
        assert not fb and n_c == 2


def _learned_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "learned", os.path.join(ROOT, "mass-preserving-level-set-method_amd", "learned.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("seed", [1001, 1002])
def test_ablation_network_gains_over_sweeps_alone(seed):
    """What the trained network adds: on held-out random shapes and flows
    (seeds the training never drew) the MPLS loop runs 200 steps, then the
    field after 1 and after 3 more advection steps is reinitialised three
    ways.  Network + 2 sweeps must land closer to the full plain
    reinitialisation than the 2 sweeps alone (band rms error over
    |phi| < 6h, in units of h)."""
    lsmp = load_lsmp()
    L = _learned_module()
    rng = np.random.default_rng(seed)
    n = 96
    h = 1.0 / n
    with L.load_net() as net, lsmp.Net() as zero:
        zero.set_params(np.zeros(zero.nparam, np.float32))
        with lsmp.Solver(lsmp.make_params(n, 3, precision=lsmp.FP32)) as s:
            s.upload_phi(L.ellipsoid_union(n, rng))
            s.reinit(1)
            s.reinit(1)
            imass, _ = s.mass_vol()
            s.set_velocity(*L.fourier_velocity(n, rng))
            for _ in range(200):
                s.rk3_advect()
                s.reinit(1)
                s.mass_fix(imass)
            for nadv in (1, 3):
                for _ in range(nadv):
                    s.rk3_advect()
                pre = s.download_phi()
                out = {}
                for kind in ("plain", "sweeps", "net"):
                    with lsmp.Solver(lsmp.make_params(n, 3, precision=lsmp.FP32)) as t:
                        t.upload_phi(pre)
                        if kind == "plain":
                            t.reinit(1)
                        else:
                            t.reinit_learned(net if kind == "net" else zero, 6.0, 2, 1e9, 1.5,
                                             lsmp.FALLBACK_RESTART)
                        out[kind] = interior(t.download_phi())
                ref = out["plain"]
                band = np.abs(ref) < 6 * h
                e_s = np.sqrt(np.mean(((out["sweeps"] - ref)[band] / h) ** 2))
                e_n = np.sqrt(np.mean(((out["net"] - ref)[band] / h) ** 2))
                say(f"seed {seed} +{nadv} steps: rms vs plain: sweeps alone {e_s:.4e}, network + sweeps {e_n:.4e}"
                    f" (ratio {e_n / e_s:.3f})")
                assert e_n < 0.97 * e_s
                s.reinit(1)
                s.mass_fix(imass)


def test_component_masses_against_scipy_labels():
    """Per-interface masses: the GPU's 6-connected components of phi > phi_min
    against scipy.ndimage.label on the same field (partition equality through
    the component of every cell's representative), and the component masses
    summing to the solver's total level-set mass."""
    ndimage = pytest.importorskip("scipy.ndimage")
    lsmp = load_lsmp()
    n = 40
    x = (np.arange(-3, n + 3) + 0.5) / n
    z, y, xx = np.meshgrid(x, x, x, indexing="ij", sparse=True)
    spheres = [((0.25, 0.3, 0.3), 0.12), ((0.7, 0.7, 0.6), 0.15), ((0.3, 0.75, 0.7), 0.1)]
    box = np.full((n + 6,) * 3, -1.0)
    for (cx, cy, cz), r in spheres:
        box = np.maximum(box, r - np.sqrt((xx - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2))
    box = np.ascontiguousarray(box)
    h = 1.0 / n
    with lsmp.Solver(lsmp.make_params(n, dim=3)) as s:
        s.upload_phi(box)
        total, _ = s.mass_vol()
        phi_min = -1.5 * h
        inner = interior(s.download_phi())
        lab, nlab = ndimage.label(inner > phi_min)   # 6-connectivity by default
        cells = np.flatnonzero(lab.ravel())
        masses, comp = s.component_masses(phi_min, cells)
        say(f"components gpu {masses.size} scipy {nlab}; masses {masses}")
        assert masses.size == nlab == 3
        # same partition: scipy label -> gpu component is a bijection
        pairs = set(zip(lab.ravel()[cells].tolist(), comp.tolist()))
        assert len(pairs) == nlab and len({p[1] for p in pairs}) == nlab and min(comp) >= 0
        assert abs(masses.sum() - total) <= 1e-12 * total
        # each component's mass equals the mass of its sphere alone
        centre_cells = [int(((round(cz * n - 0.5)) * n + round(cy * n - 0.5)) * n + round(cx * n - 0.5))
                        for (cx, cy, cz), _ in spheres]
        _, cc = s.component_masses(phi_min, centre_cells)
        assert sorted(cc.tolist()) == [0, 1, 2]
        for (c3, r), k in zip(spheres, cc):
            cx, cy, cz = c3
            one = np.ascontiguousarray(np.broadcast_to(r - np.sqrt((xx - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2),
                                                       box.shape))
            s.upload_phi(np.maximum(one, -1.0))
            m1, _ = s.mass_vol()
            assert abs(masses[k] - m1) <= 1e-12 * m1
