"""Writes tests/golden/rod64_path16.json: converged eigenvalues of the 64x64
rod unit cell (eps 8.9, r 0.2, 1/eps sampled at cell corners) at 16 Bloch
parameters arc-length uniform on Gamma-X-M-Gamma, 10 smallest non-zero
eigenvalues of K (the reference prints lambda - 1 of K + M).

Source of the numbers: the oracle's FP64 block PINVIT with the reference's
line-ordered smoother (colour_order=0, quirk=True), tolerance 1e-11, from a
seeded start block.  Two k-points (Gamma and X) are cross-checked against a
dense generalised eigensolve (scipy.linalg.eigh) of the oracle's element
assembly, and the difference is stored in the file.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np
import scipy.linalg as sl

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (k-path definition)
import oracle as O  # noqa: E402

N, N0, L, NK, NEIG, BLOCK, TOL = 64, 4, 4, 16, 10, 12, 1e-11


def main():
    eta = O.rod_eta(N)
    kx, ky = bench.kpath(NK)
    mg = O.Multigrid(N0, L, eta, colour_order=0, quirk=True)
    lam, rel, its = [], [], []
    for p in range(NK):
        mg.set_bloch(kx[p], ky[p])
        E0 = O.random_block(BLOCK, 2 * N * N, 100 + p)
        l, _, r, _, it = mg.block_solve(E0, NEIG, tol=TOL, maxit=2000)
        assert (r[:NEIG] <= TOL).all(), (p, r)
        lam.append(l[:NEIG].tolist())
        rel.append(float(r[:NEIG].max()))
        its.append(int(it))
        print("k %2d (%.4f, %.4f): %d steps, lam %s" % (p, kx[p], ky[p], it, np.array2string(l[:NEIG], precision=6)),
              flush=True)
    dense = {}
    for p in (0, 5):
        g = O.Grid2D(eta, kx=kx[p], ky=ky[p])
        w = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1), eigvals_only=True,
                    subset_by_index=[0, N * N + NEIG + 2])
        w = w[np.abs(w) > 1e-6][:NEIG]
        dense[str(p)] = float(np.max(np.abs(w - np.array(lam[p])) / w))
        print("dense check k %d: max rel diff %.3e" % (p, dense[str(p)]), flush=True)
    out = dict(workload="64x64 rod (eps 8.9, r 0.2), 16 k on Gamma-X-M-Gamma, 10 eigenvalues of K",
               n=N, n0=N0, levels=L, kx=kx.tolist(), ky=ky.tolist(), neig=NEIG, block=BLOCK, tol=TOL,
               lam=lam, max_relres=rel, oracle_steps=its, dense_check_max_rel_diff=dense)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rod64_path16.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
