"""Pins of the oracle's two-marker MPLS branch (interface_solver.f90:34-128),
CPU only.

The merge rule is pinned by what it must do to shapes: two markers apart
merge into their pointwise maximum (the union's field, no reconstruction);
overlapping markers flag a reconstruction and keep the union as the positive
set.  Which of the two positive values the overlap keeps (the smaller, :99)
is restated literally and checked only through GPU = oracle.  The
orchestration is pinned against the one-field MPLS step: while the markers are
apart, a two-marker step leaves each marker exactly as a one-field step of
that marker alone, with its own initial mass, does.
"""
import numpy as np

import oracle_lib as O


def ball_box(o, c, r):
    """r - |x - c| on the oracle's box, ghosts copied from the nearest
    interior cell (zero-gradient)."""
    p = o.p
    i = (np.clip(np.arange(o.nxb) - 3, 0, p.nx - 1) + 0.5) * p.dx
    j = (np.clip(np.arange(o.nyb) - 3, 0, p.ny - 1) + 0.5) * p.dy
    if p.dim == 3:
        k = (np.clip(np.arange(o.nzb) - 3, 0, p.nz - 1) + 0.5) * p.dz
        z, y, x = np.meshgrid(k, j, i, indexing="ij", sparse=True)
        return np.ascontiguousarray(r - np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2))
    y, x = np.meshgrid(j, i, indexing="ij", sparse=True)
    d = r - np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2)
    return np.ascontiguousarray(np.broadcast_to(d, (o.nzb, o.nyb, o.nxb)))


def mass_of(p, box):
    o = O.Oracle(p)
    o.field("PHI")[:] = box
    o.ls_mv()
    return o.get("mass")


def test_merge_of_markers_apart_is_their_maximum():
    p = O.make_params(20, dim=3)
    o = O.Oracle(p)
    m1 = ball_box(o, (0.3, 0.3, 0.3), 0.15)
    m2 = ball_box(o, (0.72, 0.7, 0.68), 0.15)
    o.field("MK1")[:] = m1
    o.field("MK2")[:] = m2
    o.field("PHI")[:] = 7.0
    assert o.combine_markers() == 0
    assert np.array_equal(o.field("PHI"), np.maximum(m1, m2))


def test_merge_of_overlapping_markers_flags_and_keeps_the_union():
    p = O.make_params(40, dim=2)
    o = O.Oracle(p)
    m1 = ball_box(o, (0.4, 0.5), 0.15)
    m2 = ball_box(o, (0.6, 0.5), 0.15)
    o.field("MK1")[:] = m1
    o.field("MK2")[:] = m2
    assert o.combine_markers() == 1
    phi = o.field("PHI")
    assert np.array_equal(phi > 0, (m1 > 0) | (m2 > 0))
    one = (m1 > 0) != (m2 > 0)
    assert np.array_equal(phi[one], np.maximum(m1, m2)[one])          # inside exactly one: that one
    out = (m1 < 0) & (m2 < 0)
    assert np.array_equal(phi[out], np.maximum(m1, m2)[out])          # outside both: the nearer
    both = (m1 > 0) & (m2 > 0)
    assert both.any() and np.all(phi[both] > 0)


def test_two_marker_step_equals_one_field_steps_while_apart():
    p = O.make_params(16, dim=3)
    o = O.Oracle(p)
    m1 = ball_box(o, (0.3, 0.3, 0.3), 0.14)
    m2 = ball_box(o, (0.7, 0.7, 0.7), 0.12)
    o.field("MK1")[:] = m1
    o.field("MK2")[:] = m2
    im1, im2 = mass_of(p, m1), mass_of(p, m2)
    o.set("imass1", im1)
    o.set("imass2", im2)
    o.set("imass", 123.0)   # the global initial mass must not enter the marker corrections
    o.set("merged", 0)
    o.mpls_step_markers(2, 3.0, 1)
    assert o.get("merged") == 0 and o.get("imass") == 123.0
    for m, im, name in ((m1, im1, "MK1"), (m2, im2, "MK2")):
        one = O.Oracle(p)
        one.field("PHI")[:] = m
        one.set("imass", im)
        one.mpls_step(2, 3.0, 1)
        assert np.array_equal(o.field(name), one.field("PHI"))
    assert np.array_equal(o.field("PHI"), np.maximum(o.field("MK1"), o.field("MK2")))


def test_two_marker_step_merges_overlapping_markers_once():
    p = O.make_params(32, dim=2)
    o = O.Oracle(p)
    o.field("MK1")[:] = ball_box(o, (0.42, 0.5), 0.12)
    o.field("MK2")[:] = ball_box(o, (0.6, 0.5), 0.12)
    o.set("imass1", mass_of(p, o.field("MK1").copy()))
    o.set("imass2", mass_of(p, o.field("MK2").copy()))
    o.set("merged", 0)
    o.mpls_step_markers(1, 8.0, 1)
    assert o.get("merged") == 1
    mk1 = o.field("MK1").copy()
    o.set("imass", o.get("mass"))
    o.mpls_step_markers(1, 8.0, 1)
    # merged: the one-field branch, markers untouched, phi's mass held
    assert np.array_equal(o.field("MK1"), mk1)
    assert abs(o.get("mass") - o.get("imass")) / o.get("imass") < 1e-12
