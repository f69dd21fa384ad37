"""The batched GPU solver against the oracle's eigenvalues at the sizes of the
large workloads (fixtures written by tests/golden/make_golden_large.py: the
oracle's FP64 block PINVIT with the reference's line-ordered smoother, a block
two wider, tolerance 1e-10, a numpy-seeded start block):

  - the 512x512 rod path scan at four of its 64 k-points (Gamma, the point next
    to Gamma, one near X, one on M-Gamma), FP64 and mixed precision, block 12;
  - the 256x256 Brillouin-zone surface's supercell at two grid points, mixed,
    block 10;
  - the adjoint workload's 256x256 random field at two of its 48 k-points
    (Hermitian operator), FP64 and mixed;
  - the 4x4 defect supercell of the 2048x2048 workload on a 512x512 grid at two
    of its 32 k-points, mixed, 20 eigenvalues with block 24.

The GPU solver runs at its production tolerance (relative residual 1e-8) from
its own seeded start blocks.  Eigenvalue errors go with the square of the
residual, so each eigenvalue must agree with the oracle's to 1e-9 of itself.
The CPU checks below pin the fixtures themselves."""
import json
import os

import numpy as np
import pytest

import bench
import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# name -> (eta, k-list, precisions, block)
CASES = {
    "rod512_path64": (lambda: O.rod_eta(512), lambda: bench.kpath(64), (0, 1), 12),
    "surface256_bz32": (lambda: bench.supercell_eta(256, seed=1), lambda: bench.bz_grid(32), (1,), 10),
    "randfield256_path48": (lambda: bench.random_field_eta(256, seed=4), lambda: bench.kpath(48), (0, 1), 12),
    "defect512_path32": (lambda: bench.defect_supercell_eta(512), lambda: bench.kpath(32), (1,), 24),
}


def load(name):
    path = os.path.join(HERE, name + ".json")
    if not os.path.exists(path):
        pytest.skip("fixture %s not written" % name)
    with open(path) as f:
        g = json.load(f)
    ref = g.get("hermitian", g.get("quirk"))
    return g, ref


@pytest.mark.parametrize("name", list(CASES))
def test_fixture_consistent(name):
    """CPU: the fixture's k-points are the workload's, its eigenvalues are
    sorted and positive, and every point reached its tolerance."""
    g, ref = load(name)
    kx, ky = CASES[name][1]()
    assert np.allclose(kx[g["sel"]], g["kx"]) and np.allclose(ky[g["sel"]], g["ky"])
    lam = np.array(ref["lam"])
    assert lam.shape == (len(g["sel"]), g["neig"])
    assert (lam > 0).all() and (np.diff(lam, axis=1) >= 0).all()
    assert all(r <= t for r, t in zip(ref["max_relres"], ref["tol"]))


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_matches_oracle_fixture(name, method):
    """method 0: block PINVIT; 1: LOBPCG (blocks of at most 16)."""
    import maxwell_amd as M
    if method == 1 and CASES[name][3] > 16:
        pytest.skip("LOBPCG takes blocks of at most 16")
    g, ref = load(name)
    eta_of, kpts, precs, block = CASES[name]
    eta = eta_of()
    kx, ky = kpts()
    sel = g["sel"]
    neig = g["neig"]
    lam_ref = np.array(ref["lam"])
    for prec in precs:
        print("%s method %d prec %d k %s" % (name, method, prec, sel), flush=True)
        S = M.Solver(eta, n0=g["n0"], levels=g["levels"], precision=prec, block=block, max_problems=len(sel),
                     method=method)
        r = S.solve(kx[sel], ky[sel], neig, tol=1e-8, maxit=400, seed=9)
        S.close()
        assert r["converged"].all(), r["iters"]
        assert (r["relres"][:, :neig] <= 1e-8).all()
        err = np.abs(r["lam"][:, :neig] - lam_ref) / lam_ref
        print("  steps %s, max relative eigenvalue error %.2e, per index %s" %
              (r["iters"].tolist(), err.max(), np.array2string(err.max(axis=0), precision=1)))
        assert err.max() <= 1e-9, err
