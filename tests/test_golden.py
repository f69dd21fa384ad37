"""The 64x64 rod fixture (tests/golden/rod64_path16.json, written by
tests/golden/make_golden.py): 16 Bloch parameters on Gamma-X-M-Gamma, 10
eigenvalues each.  CPU: the fixture is self-consistent and the oracle's
four-colour block PINVIT reproduces it.  GPU: the batched solver reproduces
it in FP64 and in mixed precision."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rod64_path16.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        return json.load(f)


def test_fixture_consistent(gold):
    assert len(gold["lam"]) == len(gold["kx"]) == 16
    lam = np.array(gold["lam"])
    assert (np.diff(lam, axis=1) >= 0).all() and (lam > 0).all()
    assert max(gold["max_relres"]) <= gold["tol"]
    # two k-points were checked against a dense generalised eigensolve
    assert max(gold["dense_check_max_rel_diff"].values()) < 1e-10
    # Gamma and the path's two Gamma ends carry the same spectrum
    assert np.abs(lam[0] - lam[-1]).max() / lam[0].max() < 1e-10


@pytest.mark.parametrize("p", [3, 10])
def test_oracle_four_colour_reproduces_fixture(gold, p):
    """The fixture was made with the reference's line-ordered smoother; the
    four-colour order converges to the same eigenvalues."""
    eta = O.rod_eta(gold["n"])
    mg = O.Multigrid(gold["n0"], gold["levels"], eta, colour_order=1)
    mg.set_bloch(gold["kx"][p], gold["ky"][p])
    lam, _, rel, _, _ = mg.block_solve(O.random_block(12, 2 * gold["n"] ** 2, 5), 10, tol=1e-10, maxit=500)
    ref = np.array(gold["lam"][p])
    assert (rel[:10] <= 1e-10).all()
    assert np.abs(lam[:10] - ref).max() / ref.max() < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_gpu_batch_reproduces_fixture(gold, prec):
    """All 16 k-points in one batch at the fixture's tolerance 1e-8 relative
    residual: eigenvalues within 1e-9 relative of the fixture."""
    import maxwell_amd as M
    eta = O.rod_eta(gold["n"])
    S = M.Solver(eta, n0=gold["n0"], levels=gold["levels"], precision=prec, block=12, max_problems=16)
    r = S.solve(gold["kx"], gold["ky"], neig=10, tol=1e-8, maxit=500, seed=3)
    S.close()
    assert r["converged"].all(), r["iters"]
    assert (r["relres"][:, :10] <= 1e-8).all()
    ref = np.array(gold["lam"])
    err = np.abs(r["lam"][:, :10] - ref) / ref
    assert err.max() < 1e-9, err.max()
