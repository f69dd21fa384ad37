"""Writes oracle eigenvalue fixtures for the large workloads, at a few of
their k-points each:

  rod512_path64.json        512x512 rod (eps 8.9, r 0.2), 8 levels, k-points
                            0, 1, 18, 50 of the 64 on Gamma-X-M-Gamma,
                            10 eigenvalues
  surface256_bz32.json      256x256 2x2 supercell of rods (radii and eps
                            seed-uniform, seed 1), points 528 and 645 of the
                            32x32 Brillouin-zone grid, 8 eigenvalues
  randfield256_path48.json  256x256 thresholded Gaussian random field (seed 4),
                            k-points 12 and 30 of the 48 on Gamma-X-M-Gamma,
                            10 eigenvalues of the Hermitian operator (with the
                            reference's seam-row eta reading the operator is
                            not Hermitian on this field and the block PINVIT
                            does not converge: relative residuals stay near
                            10 after 800 steps)
  defect512_path32.json     4x4 rod supercell with the centre rod removed at
                            512x512 (the geometry of the 2048x2048 workload on
                            a grid the oracle solves in minutes), k-points 10
                            and 20 of the 32 on Gamma-X-M-Gamma, 20 eigenvalues

Source of the numbers: the oracle's FP64 block PINVIT (every vector updated
every step, no locking) with the reference's line-ordered smoother
(colour_order=0), a block two wider than the workload's and a tight
tolerance, from a numpy-seeded start block -- a different generator, block
width and smoother order than the GPU solver uses, so the fixture is an
independent reference for its converged eigenvalues.  Ritz values converge
with the square of the residual, so the stored eigenvalues are far more
accurate than the stated residuals.

Run from the repository root:  python tests/golden/make_golden_large.py [case ...]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (workload definitions)
import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "rod512_path64": dict(workload="512x512 rod (eps 8.9, r 0.2), k-points of the 64 on Gamma-X-M-Gamma",
                          n=512, eta=lambda: O.rod_eta(512), kpts=lambda: bench.kpath(64), sel=[0, 1, 18, 50],
                          neig=10, block=14, tol=1e-10, quirk=(True,)),
    "surface256_bz32": dict(workload="256x256 2x2 rod supercell (seed 1), points of the 32x32 Brillouin-zone grid",
                            n=256, eta=lambda: bench.supercell_eta(256, seed=1), kpts=lambda: bench.bz_grid(32),
                            sel=[528, 645], neig=8, block=12, tol=1e-10, quirk=(True,)),
    "randfield256_path48": dict(workload="256x256 thresholded Gaussian random field (seed 4, sigma 6 cells, eps "
                                         "1 / 8.9), k-points of the 48 on Gamma-X-M-Gamma",
                                n=256, eta=lambda: bench.random_field_eta(256, seed=4),
                                kpts=lambda: bench.kpath(48), sel=[12, 30], neig=10, block=12, tol=1e-10,
                                quirk=(False,)),
    "defect512_path32": dict(workload="512x512 4x4 rod supercell (eps 8.9, r 0.2 of the sub-cell) with the "
                                      "centre rod removed, k-points of the 32 on Gamma-X-M-Gamma",
                             n=512, eta=lambda: bench.defect_supercell_eta(512), kpts=lambda: bench.kpath(32),
                             sel=[10, 20], neig=20, block=22, tol=1e-10, quirk=(True,)),
}
# k-points next to Gamma: the lowest band's eigenvalue is small there, and the
# FP64 rounding of K e over |lambda| floors the relative residual near 1e-9
TOL_NEAR_GAMMA = 3e-9
MAXIT = 800


def run(name):
    c = CASES[name]
    n = c["n"]
    L = int(round(np.log2(n // 4)))
    eta = c["eta"]()
    kx, ky = c["kpts"]()
    out = dict(workload=c["workload"], n=n, n0=4, levels=L, neig=c["neig"], oracle_block=c["block"],
               smoother="line-ordered (reference order)", sel=c["sel"],
               kx=[float(kx[p]) for p in c["sel"]], ky=[float(ky[p]) for p in c["sel"]])
    for quirk in c["quirk"]:
        key = "quirk" if quirk else "hermitian"
        mg = O.Multigrid(4, L, eta, colour_order=0, quirk=quirk)
        lam, rel, its, tols, secs = [], [], [], [], []
        for j, p in enumerate(c["sel"]):
            mg.set_bloch(kx[p], ky[p])
            tol = c["tol"] if np.hypot(kx[p], ky[p]) > 0.5 or (kx[p] == 0 and ky[p] == 0) else TOL_NEAR_GAMMA
            E0 = O.random_block(c["block"], 2 * n * n, 1000 + p)
            t0 = time.perf_counter()
            l, _, r, _, it = mg.block_solve(E0, c["neig"], tol=tol, maxit=MAXIT)
            dt = time.perf_counter() - t0
            assert (r[:c["neig"]] <= tol).all(), (name, key, p, r[:c["neig"]], it)
            lam.append([float("%.17g" % v) for v in l[:c["neig"]]])
            rel.append(float(r[:c["neig"]].max()))
            its.append(int(it))
            tols.append(tol)
            secs.append(round(dt, 1))
            print("%s %s k %d (%.4f, %.4f): %d steps %.0f s, relres %.2e, lam %s" %
                  (name, key, p, kx[p], ky[p], it, dt, r[:c["neig"]].max(),
                   np.array2string(l[:c["neig"]], precision=8)), flush=True)
        out[key] = dict(lam=lam, max_relres=rel, tol=tols, oracle_steps=its, seconds=secs)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(out, f, indent=1)


def main():
    names = sys.argv[1:] or list(CASES)
    for name in names:
        run(name)


if __name__ == "__main__":
    main()
