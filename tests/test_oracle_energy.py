"""The oracle's energy_and_momentum (output.f90:250-397) pinned by closed
forms: uniform and linear velocity fields, whose CCD derivatives are exact
away from the boundary, give the kinetic energy, the viscous dissipation Q
and the momentum in closed form; a sphere gives its centroid, volume and
area; a plane gives its area exactly.  The checks are chosen so that the
plausible slips fail: the 1/2 on the cross terms of Q and their pairing
(uy + vx, uz + wx, wy + vz), momentum weighted by rho instead of the marker,
the inverse density and viscosity branches, the marker > 0.99 cut.
"""
import numpy as np
import pytest

import oracle_lib as O

N = 64


def oracle(n=N, dim=3, **kw):
    o = O.Oracle(O.make_params(n, dim=dim, **kw))
    o.initial_phi(2 if dim == 3 else 1, 0)
    return o


def set_velocity(o, fu, fv, fw):
    """Cell-centre velocity on the interior, then the zero-gradient ghost fill
    the reference's velocity routine ends with (velocity.f90:40-46)."""
    for name, f in (("U", fu), ("V", fv), ("W", fw)):
        a = o.field(name)
        nzb, nyb, nxb = a.shape
        x = ((np.arange(nxb) - 2.5) / o.p.nx)[None, None, :]
        y = ((np.arange(nyb) - 2.5) / o.p.ny)[None, :, None]
        z = ((np.arange(nzb) - 2.5) / o.p.nz)[:, None, None]
        a[...] = f(x, y, z) + 0.0 * (x + y + z)
        o.bc_field(name)


def inside_volume(o):
    """Volume of the cells with H > 0.99 (the kinetic-energy and dissipation
    cells), from the oracle's own Heaviside."""
    o.ls_funs()
    return np.count_nonzero(o.interior("HEAVY") > 0.99) * o.p.dx * o.p.dy * o.p.dz


def zero(x, y, z):
    return 0.0 * x


def test_uniform_velocity():
    o = oracle()
    a, b, c = 0.5, -0.25, 0.125   # exact in binary: every derivative is exactly 0
    set_velocity(o, lambda x, y, z: a + 0 * x, lambda x, y, z: b + 0 * x, lambda x, y, z: c + 0 * x)
    e = o.energy_momentum(1.0)
    vin = inside_volume(o)
    assert abs(e["Q"]) < 1e-30
    assert (e["px"], e["py"], e["pz"]) == pytest.approx((a, b, c), rel=1e-14)
    assert e["Ke"] == pytest.approx(0.5 * (a * a + b * b + c * c) * vin, rel=1e-13)


def test_momentum_weights_the_marker_not_the_density():
    """With rho_12 = 0.5 and the inverse branch, rho is about 0.5 inside: the
    momentum stays the velocity (marker-weighted), the kinetic energy halves."""
    o = oracle(rho_12=0.5, inverse=1)
    set_velocity(o, lambda x, y, z: 0.5 + 0 * x, zero, zero)
    e = o.energy_momentum(1.0)
    vin = inside_volume(o)
    assert e["px"] == pytest.approx(0.5, rel=1e-14)
    assert e["Ke"] == pytest.approx(0.5 * 0.25 * 0.5 * vin, rel=1e-2)
    assert e["Po"] == pytest.approx(0.5 * 0.35 * vin, rel=2e-2)


S = 0.5
# (u, v, w, lq): Q = 2 mu lq V_in with lq from output.f90:350
SHEARS = [
    ((lambda x, y, z: S * x), zero, zero, S * S),                           # ux^2
    ((lambda x, y, z: S * y), zero, zero, 0.5 * S * S),                     # (uy + vx)^2 / 2
    ((lambda x, y, z: S * y), (lambda x, y, z: S * x), zero, 2 * S * S),    # the pair adds
    ((lambda x, y, z: S * y), (lambda x, y, z: -S * x), zero, 0.0),         # rotation: no dissipation
    (zero, (lambda x, y, z: S * z), (lambda x, y, z: S * y), 2 * S * S),    # wy + vz
    ((lambda x, y, z: S * z), zero, (lambda x, y, z: -S * x), 0.0),         # uz + wx
    (zero, zero, (lambda x, y, z: S * z), S * S),                           # wz^2
]


@pytest.mark.parametrize("k", range(len(SHEARS)))
def test_linear_velocity_dissipation(k):
    fu, fv, fw, lq = SHEARS[k]
    o = oracle()
    set_velocity(o, fu, fv, fw)
    e = o.energy_momentum(1.0)
    vin = inside_volume(o)
    # the clamped ghosts perturb the CCD derivatives near the walls; the
    # sphere is 13 cells from them, where they are exact to ~4e-8
    assert e["Q"] == pytest.approx(2.0 * lq * vin, rel=1e-6, abs=1e-12 * vin)


def test_viscosity_ratio_inverse():
    o = oracle(inverse=1)
    set_velocity(o, (lambda x, y, z: S * x), zero, zero)
    vin = inside_volume(o)
    e = o.energy_momentum(0.25)   # inverse: mu = H mu_12 + (1 - H), about 0.25 inside
    assert e["Q"] == pytest.approx(2.0 * 0.25 * S * S * vin, rel=1e-2)


def test_sphere_centroid_volume_area():
    o = oracle()
    set_velocity(o, zero, zero, zero)
    e = o.energy_momentum(1.0)
    r = 0.15
    assert (e["xc"], e["yc"], e["zc"]) == pytest.approx((0.35, 0.35, 0.35), abs=1e-5)
    assert e["v"] == pytest.approx(4.0 / 3.0 * np.pi * r ** 3, rel=2e-2)
    assert e["A"] == pytest.approx(4.0 * np.pi * r ** 2, rel=1e-1)
    assert e["Ke"] == 0.0 and e["Q"] == 0.0


def test_plane_area_exact():
    """phi = z0 - z: exactly one layer of cells has 0 < phi < dx, so A = 1 (the
    unit cross-section) with dx = dy = dz."""
    o = oracle(n=32)
    phi = o.field("PHI")
    nzb = phi.shape[0]
    z = ((np.arange(nzb) - 2.5) / 32)[:, None, None]
    phi[...] = 0.4037 - z + 0.0 * phi
    e = o.energy_momentum(1.0)
    assert e["A"] == pytest.approx(1.0, rel=1e-12)


def test_vof_marker_equals_heavy_marker():
    o = oracle(n=32)
    set_velocity(o, (lambda x, y, z: S * y), (lambda x, y, z: S * x), (lambda x, y, z: S * z))
    o.ls_funs()
    o.field("VOF")[...] = o.field("HEAVY")
    assert o.energy_momentum(0.7, True) == o.energy_momentum(0.7, False)


def test_2d_shear():
    """2D: w is zero and the z derivatives of the 7-plane box vanish (the
    circle is 0.1 from the top wall: 13 cells at 128^2)."""
    o = oracle(n=128, dim=2)
    set_velocity(o, (lambda x, y, z: S * y), (lambda x, y, z: S * x), zero)
    e = o.energy_momentum(1.0)
    vin = inside_volume(o)
    assert e["Q"] == pytest.approx(2.0 * 2 * S * S * vin, rel=1e-6)
    assert e["pz"] == 0.0
