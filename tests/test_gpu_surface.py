"""The batched Brillouin-zone surface: a 2 x 2 supercell of rods (radii and
eps seed-uniform), the uniform 32 x 32 Bloch grid over [-pi, pi]^2 of the
reference's blochsurface (1024 eigenproblems), 8 eigenvalues with block 10,
mixed precision, 64 problems per launch with per-problem convergence and
locking, smoother tilings from the autotuner (tuned at the first launch,
kept for the other fifteen).

Parity at 64^2 on an 8 x 8 grid: eigenvalues against dense generalised
eigensolves of the oracle's assembly.  At the full 256^2 size, properties the
mathematics fixes: lambda(k) = lambda(-k) (time reversal; the grid is
symmetric, index i <-> 31 - i), and the two ends of the grid (k = -pi and pi)
are the same Bloch problem."""
import time

import numpy as np
import pytest
import scipy.linalg as sl

import bench
import oracle as O

pytestmark = pytest.mark.gpu


def solve_surface(M, eta, kx, ky, n0, per_launch, neig, block, tol, prec, autotune=False):
    S = M.Solver(eta, n0=n0, precision=prec, block=block, max_problems=per_launch, autotune=autotune)
    lam = np.zeros((len(kx), block))
    rel = np.zeros((len(kx), block))
    conv = np.zeros(len(kx), bool)
    t0 = time.perf_counter()
    for a in range(0, len(kx), per_launch):
        r = S.solve(kx[a:a + per_launch], ky[a:a + per_launch], neig, tol=tol, maxit=400, seed=a + 1)
        lam[a:a + per_launch], rel[a:a + per_launch] = r["lam"], r["relres"]
        conv[a:a + per_launch] = r["converged"]
    dt = time.perf_counter() - t0
    tuning = S.tuning()
    S.close()
    return lam, rel, conv, dt, tuning


def test_surface_64_matches_dense():
    import maxwell_amd as M
    n, nk = 64, 8
    eta = bench.supercell_eta(n, seed=1)
    kx, ky = bench.bz_grid(nk)
    lam, rel, conv, _, _ = solve_surface(M, eta, kx, ky, 4, 16, 8, 10, 1e-9, M.MX_PREC_MIXED)
    assert conv.all() and (rel[:, :8] <= 1e-9).all()
    for p in (0, 9, 27, 63):
        g = O.Grid2D(eta, kx=kx[p], ky=ky[p])
        w = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1), eigvals_only=True,
                    subset_by_index=[0, n * n + 12])
        w = w[np.abs(w) > 1e-6][:8]
        assert np.abs(lam[p, :8] - w).max() / w.max() < 1e-9, (p, lam[p, :8], w)


def test_surface_256_full_grid_symmetries():
    import maxwell_amd as M
    n, nk = 256, 32
    eta = bench.supercell_eta(n, seed=1)
    kx, ky = bench.bz_grid(nk)
    lam, rel, conv, dt, tuning = solve_surface(M, eta, kx, ky, 4, 64, 8, 10, 1e-8, M.MX_PREC_MIXED, autotune=True)
    print("surface 256^2, 1024 k, 8 eigs: %.2f s, %.1f eigenpairs/s, tilings %s" % (dt, 1024 * 8 / dt, tuning))
    assert conv.all(), np.nonzero(~conv)
    assert (rel[:, :8] <= 1e-8).all()
    L = lam[:, :8].reshape(nk, nk, 8)  # [iy, ix]
    scale = L.max()
    # time reversal: (iy, ix) <-> (nk-1-iy, nk-1-ix)
    assert np.abs(L - L[::-1, ::-1]).max() / scale < 1e-9
    # k = -pi and k = pi are the same problem
    assert np.abs(L[:, 0] - L[:, -1]).max() / scale < 1e-9
    assert np.abs(L[0, :] - L[-1, :]).max() / scale < 1e-9
