"""Pins of the oracle's CLSVOF restatement (Vortex/src/interface_solver/
clsvof.f90:1-324; interface_solver.f90:234-244), CPU only.

* The face velocities (Vortex/src/setup/velocity.f90:20-33 and the 2D
  cases') have zero discrete divergence on a uniform grid, an identity of
  the separable fields at face midpoints.
* A full and an empty VOF field stay full and empty under any face velocity:
  the divergence term vof_old dt du/dd must cancel the full cells' fluxes
  exactly as written (a dropped term or a flipped sign leaves O(dt div u)).
* Under a constant velocity the VOF volume changes only by the fluxes
  through the domain's end faces (the face fluxes telescope: F(i) is the
  flux through face i + 1/2), and the last pass only clamps to [0, 1].
* The WLIC-THINC flux of a cell cut by a planar interface normal to the flow
  equals the integral, by quadrature, of the THINC profile
  0.5 (1 + alpha tanh(beta (xi - xc))) over the swept part of the upwind
  cell, with xc found by root finding from the cell's volume fraction: this
  pins the closed forms for xc and the log-cosh flux, the upwind choice and
  isgn, independently of the formulas.
* clsvof_recon runs the reference's sequence (phi <- 2 vof - 1, btn = 0
  reinitialisation to 5 ls_wid, the old phi kept where |phi_old| > 1.5
  ls_wid, btn = 1 reinitialisation): checked against the same sequence of the
  oracle's pinned pieces.
* A few CLSVOF steps (LeVeque field, reconstruction every step) keep
  0 <= vof <= 1, the VOF volume to 1e-3, and phi's sign with the VOF away
  from the interface.
"""
import numpy as np
from scipy.integrate import quad
from scipy.optimize import brentq

import oracle_lib as O

BETA = 2.3


def oracle(n, dim=3, **kw):
    p = O.make_params(n, dim=dim, **kw)
    return O.Oracle(p)


def test_full_and_empty_vof_stay_full_and_empty():
    for dim in (2, 3):
        for fill in (0.0, 1.0):
            o = oracle(12, dim=dim)
            o.initial_phi(2 if dim == 3 else 1, 0)          # normals from a sphere / circle
            o.velocity_faces(2 if dim == 3 else 1, 0.4, 3.0)  # non-uniform, divergence-free field
            o.field("VOF")[:] = fill
            o.field("VOF_OLD")[:] = fill
            o.vof_wlic()
            v = o.interior("VOF")
            if fill == 0.0:
                assert np.all(v == 0.0)
            else:
                assert np.max(np.abs(v - 1.0)) < 1e-13, np.max(np.abs(v - 1.0))


def mac_divergence(o):
    """(u_i - u_{i-1})/dx + (v_j - v_{j-1})/dy [+ (w_k - w_{k-1})/dz] on the
    cells 2 .. n - 1 of each direction (z, y, x)."""
    u, v, w = o.field("UF"), o.field("VF"), o.field("WF")
    p = o.p
    if p.dim == 3:
        c = np.s_[4:-4, 4:-4, 4:-4]
        return ((u[c] - u[4:-4, 4:-4, 3:-5]) / p.dx + (v[c] - v[4:-4, 3:-5, 4:-4]) / p.dy
                + (w[c] - w[3:-5, 4:-4, 4:-4]) / p.dz)
    c = np.s_[3, 4:-4, 4:-4]
    return (u[c] - u[3, 4:-4, 3:-5]) / p.dx + (v[c] - v[3, 3:-5, 4:-4]) / p.dy


def test_face_velocities_are_discretely_divergence_free():
    """Each face component is its field at the face's midpoint and the other
    coordinates at the cell centre: for these separable fields
    sin^2(pi (x + h/2)) - sin^2(pi (x - h/2)) = sin(2 pi x) sin(pi h), so the
    face differences cancel exactly (a component taken at the cell centre,
    or the face shift on the wrong coordinate, leaves O(h) or O(1)).  The
    last cell of each direction sees the ghost coordinate clamped to the
    wall, as in the reference's grid, and is left out."""
    for dim, field, n in ((3, 2, 20), (2, 1, 32), (2, 0, 24)):
        o = oracle(n, dim=dim)   # equal spacings: sin(pi h) / h is the same factor in every direction
        o.velocity_faces(field, 0.4, 3.0)
        d = mac_divergence(o)
        scale = np.max(np.abs(o.field("UF"))) / o.p.dx
        assert scale > 1.0
        assert np.max(np.abs(d)) <= 1e-13 * scale, (dim, field, np.max(np.abs(d)))


def test_constant_velocity_conserves_volume():
    """A ball of VOF away from the walls under u = 0.7 (x only): the end
    faces carry nothing, so the x pass (read back from its face fluxes) keeps
    the sum of vof to round-off; the last pass only clamps to [0, 1], which
    trims the THINC profile's overshoots (4e-13 here)."""
    o = oracle(20, dim=3)
    o.initial_phi(2, 1)                      # sphere r = 0.15 at 0.35: VOF by sub-sampling
    o.field("UF")[:] = 0.7
    o.field("VF")[:] = 0.0
    o.field("WF")[:] = 0.0
    o.field("VOF_OLD")[:] = o.field("VOF")
    old = o.field("VOF")[3:-3, 3:-3, 3:-3].copy()
    o.vof_wlic()
    F = o.field("SX1")[3:-3, 3:-3, :]         # F[..., b]: the flux through face i + 1/2, i = b - 2
    assert np.all(F[..., 2] == 0.0) and np.all(F[..., -4] == 0.0)   # the end faces 1/2 and nx + 1/2
    x_pass = old - (F[..., 3:-3] - F[..., 2:-4]) / o.p.dx
    assert abs(x_pass.sum() - old.sum()) <= 1e-14 * old.sum()
    assert np.max(np.abs(x_pass - old)) > 1e-2   # it moved
    v = o.interior("VOF")
    assert np.array_equal(v, np.clip(x_pass, 0.0, 1.0))
    assert abs(v.sum() - old.sum()) <= 1e-12 * old.sum()


def thinc_flux_by_quadrature(vof, alpha, c, upwind_is_left, dx):
    """Flux through a face swept by c = |u| dt / dx of the upwind cell, by
    integrating the THINC profile whose centre xc is set by its volume."""
    h = lambda xi, xc: 0.5 * (1.0 + alpha * np.tanh(BETA * (xi - xc)))
    xc = brentq(lambda x: quad(h, 0.0, 1.0, args=(x,), epsabs=1e-13, epsrel=1e-13)[0] - vof, -5.0, 6.0,
                xtol=1e-15)
    # flow to the right: the right end of the upwind cell, [1 - c, 1]; to the
    # left: its left end, [0, c]
    a, b = (1.0 - c, 1.0) if upwind_is_left else (0.0, c)
    return dx * quad(h, a, b, args=(xc,), epsabs=1e-15, epsrel=1e-13)[0]


def test_thinc_flux_matches_quadrature_of_the_profile():
    n = 16
    for u in (0.9, -0.9):
        for vfrac in (0.2, 0.5, 0.83):
            for full_left in (True, False):
                o = oracle(n, dim=2)
                dx, dt = o.p.dx, o.p.dt
                i0 = 8                                    # the cut cell (Fortran index)
                vof = o.field("VOF")
                nxb = o.nxb
                # columns (box index b = i + 2): full on one side of the cut cell, empty on the other
                col = np.arange(nxb) - 2
                row = np.where(col < i0, 1.0, 0.0) if full_left else np.where(col > i0, 1.0, 0.0)
                row[i0 + 2] = vfrac
                vof[:] = row[None, None, :]
                o.field("VOF_OLD")[:] = vof
                # planar phi normal to x: n = (+-1, 0, 0), so the THINC weight is 1
                x = (np.arange(nxb) - 2.5) * dx
                o.field("PHI")[:] = ((x - (i0 - 0.5) * dx) * (-1.0 if full_left else 1.0))[None, None, :]
                o.field("UF")[:] = u
                o.field("VF")[:] = 0.0
                o.vof_wlic()
                F = o.field("SX1")
                # the face the cut cell is upwind of: its right face (u > 0) or left face (u < 0)
                face = i0 if u > 0 else i0 - 1
                got = F[3, 5, face + 2]
                alpha = -1.0 if full_left else 1.0          # vof(left) < vof(right) -> +1
                want = np.sign(u) * thinc_flux_by_quadrature(vfrac, alpha, abs(u) * dt / dx, u > 0, dx)
                assert abs(got - want) <= 1e-10 * abs(want), (u, vfrac, full_left, got, want)
                # the check tells the plausible mistakes apart: the swept part taken
                # at the wrong end of the cell, or the profile's slope reversed
                wrong_end = np.sign(u) * thinc_flux_by_quadrature(vfrac, alpha, abs(u) * dt / dx, u < 0, dx)
                wrong_slope = np.sign(u) * thinc_flux_by_quadrature(vfrac, -alpha, abs(u) * dt / dx, u > 0, dx)
                assert abs(got - wrong_end) > 1e-3 * abs(want)
                if vfrac != 0.5:
                    assert abs(got - wrong_slope) > 1e-3 * abs(want)


def test_recon_is_the_reference_sequence():
    a = oracle(14, dim=3)
    a.initial_phi(2, 1)
    a.field("PHI")[:] *= 1.4                       # not a distance: the reinitialisations move it
    b = oracle(14, dim=3)
    b.field("PHI")[:] = a.field("PHI")
    b.field("VOF")[:] = a.field("VOF")
    a.clsvof_recon()
    # the same sequence from the oracle's pieces
    old = b.field("PHI").copy()
    b.field("PHI")[:] = 2.0 * b.field("VOF") - 1.0
    b.redis(0, 5.0 * b.p.ls_wid)
    keep = np.abs(old) > 1.5 * b.p.ls_wid
    b.field("PHI")[keep] = old[keep]
    b.bc_field("PHI")
    b.redis(1, -1.0)
    assert np.array_equal(a.field("PHI"), b.field("PHI"))


def test_clsvof_steps_keep_vof_bounded_and_consistent():
    o = oracle(24, dim=3)
    o.initial_phi(2, 1)
    o.store_initial()
    v0 = o.interior("VOF").sum()
    for _ in range(4):
        o.clsvof_step(2, 3.0, 1)
    v = o.interior("VOF")
    assert v.min() >= 0.0 and v.max() <= 1.0
    assert abs(v.sum() - v0) <= 1e-3 * v0
    phi = o.interior("PHI")
    decided = (v < 0.05) | (v > 0.95)
    agree = (phi[decided] > 0) == (v[decided] > 0.5)
    assert agree.mean() > 0.99
