"""GPU parity of the two-marker MPLS branch (interface_solver.f90:34-128):
the merge kernel bit for bit against the oracle (FP64), and whole two-marker
steps -- apart, and overlapping so that they merge at the first step --
against the oracle's, to the 1e-12 bar of the one-field step."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import load_lsmp
from test_gpu_parity import TOL, pair, rel_err
from test_oracle_markers import ball_box, mass_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_combine_markers_bitwise_against_oracle():
    """Mixed signs, exact zeros and ties, with consistent ghosts."""
    o, s = pair(nx=20, ny=18, nz=22, dim=3)
    rng = np.random.default_rng(11)
    boxes = []
    for c, r in (((0.4, 0.45, 0.5), 0.2), ((0.62, 0.5, 0.48), 0.18)):
        b = ball_box(o, c, r)
        inner = b[3:-3, 3:-3, 3:-3]
        inner += 0.05 * rng.normal(size=inner.shape)
        inner[rng.random(inner.shape) < 0.02] = 0.0
        boxes.append(b)
    # ties: equal values in a slab
    boxes[1][3:-3, 8:10, 3:-3] = boxes[0][3:-3, 8:10, 3:-3]
    for b in boxes:
        o.field("PHI")[:] = b
        o.bc_field("PHI")   # ghosts copied from the perturbed interior
        b[:] = o.field("PHI")
    o.field("MK1")[:] = boxes[0]
    o.field("MK2")[:] = boxes[1]
    s.set_marker(1, boxes[0])
    s.set_marker(2, boxes[1])
    rec_o = o.combine_markers()
    rec_g = s.combine_markers()
    say(f"reconstruct oracle {rec_o} gpu {rec_g}")
    assert rec_o == 1 and rec_g
    assert np.array_equal(s.download_phi(), o.field("PHI"))
    # the markers themselves come back unchanged
    assert np.array_equal(s.get_marker(1), boxes[0]) and np.array_equal(s.get_marker(2), boxes[1])
    s.close()


def test_marker_calls_out_of_order_are_refused():
    lsmp = load_lsmp()
    s = lsmp.Solver(lsmp.make_params(8, dim=3))
    with pytest.raises(lsmp.LsmpError, match="out of order"):
        s.combine_markers()          # no marker set
    with pytest.raises(lsmp.LsmpError, match="out of order"):
        s.marker_load(1)
    with pytest.raises(lsmp.LsmpError, match="invalid"):
        s.set_marker(3, np.zeros(s.shape))
    s.set_marker(1, np.zeros(s.shape) - 1.0)
    with pytest.raises(lsmp.LsmpError, match="out of order"):
        s.combine_markers()          # marker 2 still missing
    s.close()


@pytest.mark.parametrize("case", [
    dict(nx=24, dim=3, centres=((0.3, 0.3, 0.3), (0.7, 0.7, 0.68)), radii=(0.14, 0.12), field=2, t2s=3.0),
    dict(nx=48, dim=2, centres=((0.42, 0.5), (0.6, 0.5)), radii=(0.12, 0.12), field=1, t2s=8.0),
], ids=["apart_3d", "merging_2d"])
def test_two_marker_steps_parity(case):
    o, s = pair(nx=case["nx"], dim=case["dim"])
    say("two-marker steps", {k: v for k, v in case.items()})
    m = [ball_box(o, c, r) for c, r in zip(case["centres"], case["radii"])]
    im = [mass_of(o.p, b) for b in m]
    o.field("MK1")[:] = m[0]
    o.field("MK2")[:] = m[1]
    o.set("imass1", im[0])
    o.set("imass2", im[1])
    o.set("merged", 0)
    s.set_marker(1, m[0])
    s.set_marker(2, m[1])
    # the global initial mass, used once the markers have merged, is that of
    # the initial field (problem_init.f90:140-142): here the markers' maximum
    phi0 = np.maximum(m[0], m[1])
    imass = mass_of(o.p, phi0)
    o.set("imass", imass)
    s.upload_phi(phi0)
    merged = False
    t = 0.0
    for step in range(3):
        t = t + o.p.dt
        o.mpls_step_markers(case["field"], case["t2s"], 1)
        merged, npseudo, nmass = s.mpls_step_markers(t, case["t2s"], case["field"], imass, im, merged)
        e = rel_err(s.download_phi(), o.field("PHI"))
        e1 = rel_err(s.get_marker(1), o.field("MK1"))
        e2 = rel_err(s.get_marker(2), o.field("MK2"))
        say(f"  step {step + 1}: merged {merged}/{int(o.get('merged'))}  phi {e:.3e}  markers {e1:.3e} {e2:.3e}"
            f"  pseudo {npseudo}  mass iters {nmass}")
        assert merged == bool(o.get("merged"))
        assert e <= TOL and e1 <= TOL and e2 <= TOL
    if case["dim"] == 2:
        assert merged, "the overlapping markers merge"
    else:
        assert not merged
    s.close()
