"""Host logic of the step compositions in lsmp.py, without a GPU: the C entry
points each step calls, in order, recorded through a stand-in for the
library.  The orders are the reference's:

* MPLS merged branch (main.f90:23-35, interface_solver.f90:123-129):
  velocity at t^{n+1}, RK3 advection, reinitialisation when iter mod t2red
  is 0, mass correction;
* CLSVOF merged branch (main.f90:23-27, interface_solver.f90:234-244):
  node and face velocities at t^{n+1}, vof%old <- vof (p%switch), WLIC
  advection of the VOF, RK3 advection of phi, reconstruction when iter mod
  t2red is 0.
"""
import ctypes as C

import pytest

from conftest import load_lsmp


class Recorder:
    """Every attribute is an entry point that logs its name and succeeds."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def call(*args):
            self.calls.append(name)
            return 0
        return call


def solver(dim=3):
    lsmp = load_lsmp()
    s = object.__new__(lsmp.Solver)
    s.lib = Recorder()
    s.params = lsmp.make_params(16, dim=dim)
    s.h = C.c_void_p()
    return lsmp, s


@pytest.mark.parametrize("iteration, t2red, recon", [(1, 1, True), (3, 2, False), (4, 2, True)])
def test_clsvof_step_order(iteration, t2red, recon):
    lsmp, s = solver()
    s.clsvof_step(0.1, 3.0, lsmp.VEL_LEVEQUE3D, t2red=t2red, iteration=iteration)
    want = ["lsmp_set_velocity_analytic", "lsmp_set_face_velocity_analytic", "lsmp_store_vof_old",
            "lsmp_vof_wlic", "lsmp_rk3_advect"] + (["lsmp_clsvof_recon"] if recon else [])
    assert s.lib.calls == want


@pytest.mark.parametrize("iteration, t2red, reinit", [(1, 1, True), (3, 2, False)])
def test_mpls_step_order(iteration, t2red, reinit):
    lsmp, s = solver()
    s.mpls_step(0.1, 3.0, lsmp.VEL_LEVEQUE3D, 1.0, t2red=t2red, iteration=iteration)
    want = ["lsmp_set_velocity_analytic", "lsmp_rk3_advect"] + (["lsmp_reinit"] if reinit else []) \
        + ["lsmp_mass_fix"]
    assert s.lib.calls == want
