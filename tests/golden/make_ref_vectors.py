"""Write golden vectors from the reference's own Fortran units.

Runs the reference's solve_tridiagonal (Vortex/src/others/TDMA.f90:1-36), its
CCD line solver (Vortex/lib/solver/ccd_solvers.f90:382-427 with
matrix_solver.f90:6-78) and its upwind CCD solver ("uccd", ccd_solvers.f90:
165-214, 312-345, 428-438), compiled from the reference sources by
`make -C oracle ref`, on seeded inputs, and stores inputs and outputs in
ref_vectors.npz next to this script.

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
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
