"""Write golden vectors from the reference's own Fortran units.

Runs the reference's solve_tridiagonal (Vortex/src/others/TDMA.f90:1-36), its
CCD line solver (Vortex/lib/solver/ccd_solvers.f90:382-427 with
matrix_solver.f90:6-78), its upwind CCD solver ("uccd", ccd_solvers.f90:
165-214, 312-345, 428-438), its CRWENO5 line routine with the WENO-JS face
closures (Vortex/src/interface_solver/level_set_rk3_solver.f90:288-414) and
its WENO-HJ term phyn (level_set_rk3_redis.f90:456-471), compiled from the
reference sources by `make -C oracle ref`, on seeded inputs, and stores
inputs and outputs in ref_vectors.npz next to this script.

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref_pieces.so is missing: run `make -C oracle ref`")
    rng = np.random.default_rng(20261202)
    out = {}
    # Thomas: diagonally dominant systems and CRWENO-shaped rows (A+B+C = 1)
    for tag, n in (("tdma_dd", 37), ("tdma_crw", 63)):
        if tag == "tdma_dd":
            A = rng.uniform(-0.5, 0.5, n); C = rng.uniform(-0.5, 0.5, n)
            B = np.abs(A) + np.abs(C) + rng.uniform(0.2, 1.0, n)
        else:
            w = rng.dirichlet([1.0, 1.0, 1.0], n)
            A = (2 * w[:, 0] + w[:, 1]) / 3; B = (w[:, 0] + 2 * (w[:, 1] + w[:, 2])) / 3; C = w[:, 2] / 3
        S = rng.normal(size=n)
        out[f"{tag}_A"], out[f"{tag}_B"], out[f"{tag}_C"], out[f"{tag}_S"] = A, B, C, S
        a, b, c, s = A.copy(), B.copy(), C.copy(), S.copy()
        X = np.zeros(n)
        R.ref_solve_tridiagonal(*[O._dp(v) for v in (a, b, c, s, X)], n)
        out[f"{tag}_X"] = X
    # CCD: a 2D z line (7 points), a short and a 256^3-sized x line
    for n in (7, 38, 262):
        dx = 1.0 / (n - 6)
        x = (np.arange(n) - 2.5) * dx
        f = np.sin(2 * np.pi * x) + 0.1 * rng.normal(size=n)
        fx = np.zeros(n); fxx = np.zeros(n)
        R.ref_ccd_line(n, dx, O._dp(f), O._dp(fx), O._dp(fxx))
        out[f"ccd{n}_f"], out[f"ccd{n}_dx"], out[f"ccd{n}_fx"], out[f"ccd{n}_fxx"] = f, dx, fx, fxx
        tab = np.zeros(12 * n)
        R.ref_ccd_tables(n, dx, O._dp(tab))
        out[f"ccd{n}_tables"] = tab
    # UCCD: velocity of both signs with exact zeros (the downwind branch)
    for n in (7, 38, 262):
        dx = 1.0 / (n - 6)
        x = (np.arange(n) - 2.5) * dx
        f = np.cos(2 * np.pi * x) + 0.1 * rng.normal(size=n)
        u = rng.normal(size=n)
        u[::5] = 0.0
        fx = np.zeros(n)
        R.ref_uccd_line(n, dx, O._dp(f), O._dp(u), O._dp(fx))
        out[f"uccd{n}_f"], out[f"uccd{n}_u"], out[f"uccd{n}_dx"], out[f"uccd{n}_fx"] = f, u, dx, fx
    # CRWENO5: split fluxes of smooth, discontinuous and noisy lines; odd,
    # 2-unknown and 256^3-sized lines
    for tag, n in (("smooth", 64), ("step", 63), ("noise", 255), ("tiny", 3)):
        x = (np.arange(n + 6) - 2.5) / n
        if tag == "smooth":
            u = 1.3 * np.sin(2 * np.pi * x) * np.cos(np.pi * x)
        elif tag == "step":
            u = np.where(x > 0.37, 1.0, -0.4) + 0.01 * rng.normal(size=n + 6)
        else:
            u = rng.normal(size=n + 6)
        ph = 0.15 - np.abs(x - 0.45) + 0.05 * rng.normal(size=n + 6)
        g = 0.5 * (u + np.abs(u)) * ph
        f = 0.5 * (u - np.abs(u)) * ph
        fp, gm = rng.normal(size=n + 6), rng.normal(size=n + 6)
        out[f"crw_{tag}_f"], out[f"crw_{tag}_g"] = f, g
        out[f"crw_{tag}_fp_in"], out[f"crw_{tag}_gm_in"] = fp.copy(), gm.copy()
        R.ref_crweno_flux_split(O._dp(f.copy()), O._dp(g.copy()), O._dp(fp), O._dp(gm), n)
        out[f"crw_{tag}_fp"], out[f"crw_{tag}_gm"] = fp, gm
    # phyn on second differences over six decades of scale
    v = rng.normal(size=(4096, 4)) * rng.choice([1e-7, 1e-4, 1e-2, 1.0, 1e2], size=(4096, 1))
    out["phyn_in"] = v
    out["phyn_out"] = np.array([R.ref_phyn(*r) for r in v])
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
