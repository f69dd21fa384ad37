"""Error paths and end states of the batched solver (mx_solver_create,
mx_solve, mx_get_vectors): the status codes of include/maxwell_amd.h for bad
arguments, non-finite data and a Ritz Gram matrix that is not definite; the
iteration limits (maxit = 0, an unreachable tolerance); and the Ritz vectors
a solve leaves behind for every problem of a batch whose problems finish on
steps of both parities.

The reference aborts through assert() on the failures that have a status here
(exp_bloch2d.c:1002 asserts that zhegv succeeds); the batched entry points
return the status instead and the solver stays usable."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EINVAL, ELAPACK, ENOTFINITE = 1, 4, 5


@pytest.fixture(scope="module")
def M():
    import maxwell_amd
    return maxwell_amd


def status_of(fn):
    try:
        fn()
    except RuntimeError as e:
        return int(str(e).split("error ")[1].split()[0])
    return 0


def test_create_rejects_bad_configs(M):
    eta = O.rod_eta(16)
    assert status_of(lambda: M.Solver(eta, n0=4, levels=2, gradient_steps=0)) == EINVAL
    assert status_of(lambda: M.Solver(eta, n0=4, levels=2, block=33)) == EINVAL
    assert status_of(lambda: M.Solver(eta, n0=4, levels=2, prec_steps=0)) == EINVAL
    assert status_of(lambda: M.Solver(eta, n0=4, levels=2, precision=7)) == EINVAL
    bad = eta.copy()
    bad[3, 5] = np.nan
    assert status_of(lambda: M.Solver(bad, n0=4, levels=2)) == EINVAL
    bad = eta.copy()
    bad[0, 0] = -1.0
    assert status_of(lambda: M.Solver(bad, n0=4, levels=2)) == EINVAL
    bad = eta.copy()
    bad[7, 1] = np.inf
    assert status_of(lambda: M.Solver(bad, n0=4, levels=2)) == EINVAL


def test_solve_rejects_bad_batches_and_stays_usable(M):
    eta = O.rod_eta(16)
    S = M.Solver(eta, n0=4, levels=2, block=8, max_problems=2)
    assert status_of(lambda: S.solve([0.1, 0.2, 0.3], [0.0, 0.0, 0.0], neig=6)) == EINVAL  # nprob > max
    assert status_of(lambda: S.solve([0.1], [0.0], neig=9)) == EINVAL  # neig > block
    assert status_of(lambda: S.solve([0.1], [0.0], neig=0)) == EINVAL
    assert status_of(lambda: S.solve([0.1], [0.0], neig=6, maxit=-1)) == EINVAL
    assert status_of(lambda: S.solve([0.1], [0.0], neig=6, tol=-1.0)) == EINVAL
    assert status_of(lambda: S.solve([0.1], [0.0], neig=6, tol=np.nan)) == EINVAL
    assert status_of(lambda: S.solve([np.nan], [0.0], neig=6)) == EINVAL
    r = S.solve([0.3, 0.0], [-1.1, 0.0], neig=6, tol=1e-9, maxit=300, seed=2)
    assert r["converged"].all()
    S.close()


def test_maxit_zero_returns_start_ritz_values(M):
    """maxit = 0: the start block's Ritz values and residuals, no block step."""
    eta = O.rod_eta(16)
    S = M.Solver(eta, n0=4, levels=2, block=8, max_problems=2)
    r = S.solve([0.3, 0.0], [-1.1, 0.0], neig=6, tol=1e-9, maxit=0, seed=2)
    assert (r["iters"] == 0).all() and not r["converged"].any()
    assert np.isfinite(r["lam"]).all() and np.isfinite(r["relres"]).all()
    assert (r["relres"][:, :6] > 1e-9).all()
    assert S.stats()["block_steps"] == 0
    S.close()


def test_unreachable_tolerance_stops_at_maxit(M):
    eta = O.rod_eta(16)
    S = M.Solver(eta, n0=4, levels=2, block=8, max_problems=2)
    r0 = S.solve([0.3, 0.0], [-1.1, 0.0], neig=6, tol=0.0, maxit=0, seed=2)
    r = S.solve([0.3, 0.0], [-1.1, 0.0], neig=6, tol=0.0, maxit=5, seed=2)
    assert (r["iters"] == 5).all() and not r["converged"].any()
    # five steps from the same start block have made progress
    assert np.isfinite(r["lam"]).all() and (r["relres"][:, :6].max(axis=1) < 0.5 * r0["relres"][:, :6].max(axis=1)).all()
    assert S.stats()["block_steps"] == 10
    S.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_dependent_start_vectors_give_lapack_status(M, prec):
    """A repeated start vector (or a zero one) makes the Ritz Gram matrix
    U^H M U singular to FP64 precision: MX_ELAPACK, and the next solve on the
    same solver works."""
    n = 16
    eta = O.rod_eta(n)
    S = M.Solver(eta, n0=4, levels=2, precision=prec, block=8, max_problems=1)
    E0 = O.random_block(8, 2 * n * n, 3)
    E0[5] = E0[2]
    assert status_of(lambda: S.solve([0.3], [-1.1], neig=6, E0=E0)) == ELAPACK
    E0 = O.random_block(8, 2 * n * n, 3)
    E0[7] = 0.0
    assert status_of(lambda: S.solve([0.3], [-1.1], neig=6, E0=E0)) == ELAPACK
    r = S.solve([0.3], [-1.1], neig=6, tol=1e-9, maxit=300, seed=2)
    assert r["converged"].all()
    S.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_non_finite_start_block_gives_notfinite_status(M, prec):
    n = 16
    eta = O.rod_eta(n)
    S = M.Solver(eta, n0=4, levels=2, precision=prec, block=8, max_problems=2)
    E0 = O.random_block(16, 2 * n * n, 3)
    E0[11, 17] = np.nan
    assert status_of(lambda: S.solve([0.3, 1.0], [-1.1, 0.2], neig=6, E0=E0)) == ENOTFINITE
    r = S.solve([0.3, 1.0], [-1.1, 0.2], neig=6, tol=1e-9, maxit=300, seed=2)
    assert r["converged"].all()
    S.close()


@pytest.mark.parametrize("lock", [True, False])
@pytest.mark.parametrize("prec", [0, 1])
def test_vectors_of_every_problem_are_its_ritz_vectors(M, prec, lock):
    """Problems leave the batch on steps of both parities (the solver's two
    block buffers alternate every step).  Afterwards every problem's vectors
    must be its final Ritz vectors: M-orthonormal, with the dense residual
    that mx_solve reported for them."""
    n = 32
    eta = O.rod_eta(n)
    ks = [(0.0, 0.0), (0.3, -1.1), (np.pi, np.pi), (1.0, 0.5), (np.pi, 0.0), (0.4, 2.0), (2.2, -0.7),
          (-1.3, 0.9)]
    S = M.Solver(eta, n0=4, levels=3, precision=prec, block=12, max_problems=len(ks), lock=lock)
    r = S.solve([k[0] for k in ks], [k[1] for k in ks], neig=10, tol=1e-9, maxit=300, seed=8)
    assert r["converged"].all()
    assert len(set(int(i) % 2 for i in r["iters"])) == 2, r["iters"]
    for p, (kx, ky) in enumerate(ks):
        g = O.Grid2D(eta, kx=kx, ky=ky)
        K, Mm = O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1)
        E = S.vectors(p)
        G = E.conj() @ Mm @ E.T
        assert np.abs(G - np.eye(12)).max() < 1e-9, (p, r["iters"][p])
        for i in range(10):
            e = E[i]
            me = Mm @ e
            rel = np.linalg.norm(K @ e - r["lam"][p, i] * me) / (abs(r["lam"][p, i]) * np.linalg.norm(me))
            assert rel <= 1e-9 and abs(rel - r["relres"][p, i]) <= 1e-3 * r["relres"][p, i] + 1e-13, \
                (p, i, rel, r["relres"][p, i])
    S.close()
