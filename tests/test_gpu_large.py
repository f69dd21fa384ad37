"""The large single-cell configuration: a 2048 x 2048 grid (8,388,608 edge
DOFs, 10 multigrid levels), a 4 x 4 supercell of rods (eps 8.9) with one rod
removed (a point defect), 20 eigenvalues with block 24 (Gram products over 24
vectors), mixed precision with FP64 Ritz.  Four k-points of the 32 on
Gamma-X-M-Gamma run as one batch (the path's two Gamma ends among them, which
must give the same eigenvalues); the run reports the device footprint against
the 288 GB of one MI355X and HBM throughput of the hot kernels at this size."""
import time

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


def test_defect_supercell_2048():
    import maxwell_amd as M
    n = 2048
    eta = bench.defect_supercell_eta(n)
    kx, ky = bench.kpath(32)
    sel = [0, 10, 20, 31]
    S = M.Solver(eta, n0=4, precision=M.MX_PREC_MIXED, block=24, max_problems=len(sel))
    t0 = time.perf_counter()
    r = S.solve(kx[sel], ky[sel], 20, tol=1e-8, maxit=300, seed=5)
    dt = time.perf_counter() - t0
    st = S.stats()
    print("2048^2, 4 k, 20 eigs, block 24: %.1f s (%.2f eigenpairs/s), steps %s, device %.1f GB of 288" %
          (dt, 4 * 20 / dt, r["iters"].tolist(), st["device_bytes"] / 1e9))
    for kind, name in ((10, "afw march x2"), (11, "node march x2"), (9, "gram"), (1, "edge V-cycle"),
                       (2, "node V-cycle")):
        ms, by = S.time_kernel(kind, 3)
        print("  %-14s %8.2f ms %7.0f GB/s" % (name, ms, by / ms / 1e6))
    S.close()
    assert r["converged"].all(), r["iters"]
    assert (r["relres"][:, :20] <= 1e-8).all()
    lam = r["lam"][:, :20]
    assert (lam > 0).all() and (np.diff(lam, axis=1) >= -1e-9 * lam[:, 1:]).all()
    # both Gamma ends of the path
    assert np.abs(lam[0] - lam[3]).max() / lam.max() < 1e-9
    assert st["device_bytes"] < 288e9
