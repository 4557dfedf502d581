"""The 3D LeVeque deformation case with CLSVOF (the reference's method 3) on
the GPU: sphere r = 0.15 at (0.35, 0.35, 0.35), its VOF from the reference's
30^3 sub-samples per cell (Vortex/src/setup/problem_init.f90:32-66), face and
node velocities of Vortex/src/setup/velocity.f90 with t2s = T, dt = rdt =
0.1 dx, ls_wid = 1.5 dx, and the merged clsvof branch every step
(interface_solver.f90:234-244: WLIC-THINC VOF advection, CRWENO5 RK3 of phi,
reconstruction every t2red steps).

Runs `steps` steps (default: to T) and prints one JSON line: ms/step, the
VOF volume change, the level-set and VOF shape errors against the initial
fields when the run reaches T, and with --profile the per-kernel times of one
more step.  The reference's own Vortex deck is --n 64 --cfl 0.5 --t2red -1.

    python mass-preserving-level-set-method_amd/clsvof.py --n 256 --steps 20 --profile
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
import lsmp  # noqa: E402
from leveque import sphere_box  # noqa: E402

UG = 30
KERNELS = ("velocity", "face_velocity", "norm_x", "norm_y", "norm_z", "wlic", "adv_x", "adv_y", "adv_z",
           "recon_start", "sign_init", "grad_max", "scale_max", "redis_stage", "recon_keep")


def sphere_vof_box(n):
    """The reference's VOF: for each cell, 1/30^3 added once per sub-sample
    (x_face + ii dx / 30, ii = 1..30 per axis) inside the sphere
    (problem_init.f90:41-59).  The sum depends only on the count of inside
    samples, so it is the sequential partial sum for that count.  Cells far
    from the sphere's surface are counted without sampling."""
    dx = 1.0 / n
    c, r = float(np.float32(0.35)), float(np.float32(0.15))
    # x(i) = (i - 0.5) dx with i - 0.5 in single precision, x(0) the wall
    # (Vortex/lib/data/tree.f90:254-256); sample origin 0.5 (x(i) + x(i-1))
    xi = np.concatenate(([0.0], (np.arange(1, n + 1, dtype=np.float32) - np.float32(0.5)).astype(np.float64) * dx))
    xc = xi[1:]
    face = 0.5 * (xi[1:] + xi[:-1])
    sub = np.arange(1, UG + 1) * dx / UG
    partial = np.cumsum(np.full(UG ** 3, 1.0 / float(UG) ** 3.0))
    partial = np.concatenate(([0.0], partial))
    d = np.sqrt((xc[:, None, None] - c) ** 2 + (xc[None, :, None] - c) ** 2 + (xc[None, None, :] - c) ** 2)
    vof = np.where(d < r, partial[-1], 0.0)   # (k, j, i) = (z, y, x)
    cand = np.argwhere(np.abs(d - r) < 1.5 * dx)   # cells the sphere's surface may cut
    pts = face[:, None] + sub[None, :]        # sample coordinates per cell index and ii
    for a in range(0, len(cand), 256):
        blk = cand[a:a + 256]
        zs, ys, xs = pts[blk[:, 0]], pts[blk[:, 1]], pts[blk[:, 2]]
        dist = np.sqrt((xs[:, None, None, :] - c) ** 2 + (ys[:, None, :, None] - c) ** 2
                       + (zs[:, :, None, None] - c) ** 2)
        cnt = np.count_nonzero(dist <= r, axis=(1, 2, 3))
        vof[blk[:, 0], blk[:, 1], blk[:, 2]] = partial[cnt]
    idx = np.clip(np.arange(n + 6) - 3, 0, n - 1)
    return np.ascontiguousarray(vof[np.ix_(idx, idx, idx)])


def run(n, T=3.0, steps=None, t2red=1, precision=lsmp.FP64, profile=False, cfl=0.1):
    s = lsmp.Solver(lsmp.make_params(n, dim=3, precision=precision, dt_cfl=cfl))
    s.upload_phi(sphere_box(n))
    s.store_phi0()
    s.upload_vof(sphere_vof_box(n))
    s.store_vof0()
    _, v0 = s.vof_mass_vol()
    dt = s.params.dt
    full = int(round(T / dt))
    nsteps = full if steps is None else steps
    if t2red <= 0:
        t2red = int(T / dt) + 1   # the deck's -1: no reconstruction before T (tree.f90:352)
    t = 0.0
    s.synchronize()
    t0 = time.perf_counter()
    for it in range(1, nsteps + 1):
        if t + dt > T:
            dt = T - t + 1.0e-10   # main.f90:19-21
        t += dt
        s.clsvof_step(t, T, lsmp.VEL_LEVEQUE3D, t2red=t2red, iteration=it, dt=dt)
    s.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / nsteps
    _, v1 = s.vof_mass_vol()
    l1 = (s.interface_l1(), s.vof_interface_l1()) if nsteps >= full else None
    out = {"case": f"LeVeque 3D CLSVOF {n}^3", "precision": "fp32" if precision else "fp64", "cfl": cfl,
           "steps": nsteps, "t": t, "t2red": t2red, "ms_per_step": round(ms, 3), "vof_volume": v1,
           "vof_volume_rel_change": abs(v1 - v0) / v0}
    if l1 is not None:
        out["l1"], out["vof_l1"] = l1
    if profile:
        # one more step with per-kernel timing (not in ms_per_step)
        s.set_profiling(True)
        s.clsvof_step(t + dt, T + dt, lsmp.VEL_LEVEQUE3D, t2red=t2red, iteration=nsteps + 1, dt=dt)
        s.synchronize()
        out["kernels_one_step"] = {k: dict(zip(("launches", "ms"), s.kernel_time(k))) for k in KERNELS
                                   if s.kernel_time(k)[0]}
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--T", type=float, default=3.0)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--t2red", type=int, default=1, help="reconstruction interval; 0 or less: none before T")
    ap.add_argument("--cfl", type=float, default=0.1, help="advection dt in units of dx (the Vortex deck: 0.5)")
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--profile", action="store_true", help="per-kernel times of one more step")
    a = ap.parse_args()
    print(json.dumps(run(a.n, a.T, a.steps, a.t2red, lsmp.FP32 if a.fp32 else lsmp.FP64, a.profile, a.cfl)))


if __name__ == "__main__":
    main()
