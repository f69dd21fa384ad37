"""Rayleigh-Ritz extrapolation of start blocks on the device (mx_extrapolate:
rayleigh_extrapolation, exp_blochsurf.c:332-406; exp_bloch2d.c:825-894)
against the oracle's restatement (oracle.rayleigh_extrapolation: Householder
M-orthonormalisation, Ritz matrix of K + M, Hermitian eigensolve), and its
use along a Bloch path: solves started from blocks extrapolated from the
converged blocks of the previous path points reach the same eigenvalues as
solves from seeded random starts.

As in the reference, the basis is not cleared of gradient fields at the new
Bloch parameter: the eigenvectors of other k are not divergence-free at this
one, so the Ritz step also returns near-gradient vectors with small spurious
Ritz values (16^2 rod, k2 from k1, k0: 0.2295, 0.3049, 0.7432 and 18.96 among
0.2306, 12.1083, 14.2806, ... exact), which the solve's gradient lifting then
removes; the test checks that the genuine eigenvalues are present."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import maxwell_amd
    return maxwell_amd


@pytest.mark.parametrize("sources,ieigs,block", [(2, 4, 6), (3, 4, 8), (3, 16, 12)])
def test_extrapolation_matches_oracle(M, sources, ieigs, block):
    n = 16
    eta = O.rod_eta(n)
    ks = [(0.7, -0.4), (np.pi, 1.3)]
    r = np.random.default_rng(sources * 100 + ieigs)
    dim = sources * ieigs
    H = r.standard_normal((len(ks), dim, 2 * n * n)) + 1j * r.standard_normal((len(ks), dim, 2 * n * n))
    S = M.Solver(eta, n0=4, levels=2, precision=0, block=block, max_problems=len(ks))
    E0, lam = S.extrapolate([k[0] for k in ks], [k[1] for k in ks], H, sources)
    S.close()
    for p, (kx, ky) in enumerate(ks):
        g = O.Grid2D(eta, kx=kx, ky=ky)
        e_ref, lam_ref = O.rayleigh_extrapolation(g, H[p], block)
        # Ritz values of K (the solver's convention) against K + M minus 1
        err = np.abs(lam[p] - (lam_ref - 1.0)) / np.abs(lam_ref).max()
        print("p %d Ritz values max rel err %.2e" % (p, err.max()))
        assert err.max() < 1e-12, err
        Mm = O.dense_edge(g, 0, 1)
        G = E0[p].conj() @ Mm @ E0[p].T
        assert np.abs(G - np.eye(block)).max() < 1e-11
        # each vector equals the oracle's up to a unit phase where its Ritz
        # value is separated from its neighbours
        ov = np.abs(np.einsum("ij,jk,ik->i", E0[p].conj(), Mm, e_ref))
        gap = np.minimum(np.abs(np.diff(lam_ref, prepend=-np.inf)), np.abs(np.diff(lam_ref, append=np.inf)))
        sep = gap > 1e-6 * np.abs(lam_ref).max()
        assert sep.sum() >= block // 2
        assert np.abs(ov[sep] - 1.0).max() < 1e-9, ov


def test_extrapolation_rejects_bad_arguments(M):
    n = 16
    S = M.Solver(O.rod_eta(n), n0=4, levels=2, precision=0, block=8, max_problems=2)
    H = np.zeros((1, 6, 2 * n * n), np.complex128)
    with pytest.raises(RuntimeError, match="error 1"):
        S.extrapolate([0.3], [0.2], H, 2)  # 6 basis vectors for a block of 8
    H = np.ones((1, 50, 2 * n * n), np.complex128)
    with pytest.raises(RuntimeError, match="error 1"):
        S.extrapolate([0.3], [0.2], H, 2)  # more than 48
    H = np.ones((3, 8, 2 * n * n), np.complex128)
    with pytest.raises(RuntimeError, match="error 1"):
        S.extrapolate([0.3, 0.1, 0.2], [0.2, 0.0, 0.1], H, 1)  # more problems than max_problems
    # a NaN in the basis: the Ritz matrix is not finite
    r = np.random.default_rng(1)
    H = r.standard_normal((1, 8, 2 * n * n)) + 0j
    H[0, 3, 17] = np.nan
    with pytest.raises(RuntimeError, match="error 5"):
        S.extrapolate([0.3], [0.2], H, 1)
    S.close()
    # a grid too small for the Householder units: 8 x 8 has 8 * 4 = 32
    S = M.Solver(O.rod_eta(8), n0=4, levels=1, precision=0, block=8, max_problems=1)
    with pytest.raises(RuntimeError, match="error 1"):
        S.extrapolate([0.3], [0.2], np.ones((1, 36, 128), np.complex128), 3)
    S.close()


def test_single_source_is_rayleigh_ritz(M):
    """One source block: the Ritz pairs of K + M over its span (the M-Gram's
    generalised eigenproblem gives the same values)."""
    n, m = 16, 6
    eta = O.rod_eta(n)
    r = np.random.default_rng(4)
    H = r.standard_normal((1, 10, 2 * n * n)) + 1j * r.standard_normal((1, 10, 2 * n * n))
    S = M.Solver(eta, n0=4, levels=2, precision=0, block=m, max_problems=1)
    E0, lam = S.extrapolate([1.1], [-0.6], H, 1)
    S.close()
    g = O.Grid2D(eta, kx=1.1, ky=-0.6)
    K, Mm = O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1)
    B = H[0]
    w = sl.eigh(B.conj() @ (K + Mm) @ B.T, B.conj() @ Mm @ B.T, eigvals_only=True)
    assert np.abs(lam[0] - (w[:m] - 1.0)).max() < 1e-11 * w[m - 1]


@pytest.mark.parametrize("prec", [0, 1])
def test_path_continuation_with_extrapolated_starts(M, prec):
    """The reference's path loop (exp_bloch2d.c:825-894) in batches: solve
    two path points from random starts, then start the third from the
    first-order extrapolation of their blocks and the fourth from the
    second-order extrapolation of all three."""
    n = 32
    eta = O.rod_eta(n)
    kp = [(0.40 + 0.15 * i, 0.05 * i) for i in range(4)]
    S = M.Solver(eta, n0=4, levels=3, precision=prec, block=12, max_problems=2)
    r01 = S.solve([kp[0][0], kp[1][0]], [kp[0][1], kp[1][1]], neig=10, tol=1e-9, maxit=300, seed=3)
    assert r01["converged"].all()
    E_0, E_1 = S.vectors(0), S.vectors(1)
    E0_2, lam_x = S.extrapolate([kp[2][0]], [kp[2][1]], np.vstack([E_1, E_0])[None], 2)
    rx = S.solve([kp[2][0]], [kp[2][1]], neig=10, tol=1e-9, maxit=300, E0=E0_2)
    E_2 = S.vectors(0)
    rr = S.solve([kp[2][0]], [kp[2][1]], neig=10, tol=1e-9, maxit=300, seed=3)
    E0_3, _ = S.extrapolate([kp[3][0]], [kp[3][1]], np.vstack([E_2, E_1, E_0])[None], 3)
    rx3 = S.solve([kp[3][0]], [kp[3][1]], neig=10, tol=1e-9, maxit=300, E0=E0_3)
    rr3 = S.solve([kp[3][0]], [kp[3][1]], neig=10, tol=1e-9, maxit=300, seed=3)
    S.close()
    print("prec %d steps: random %s / %s, extrapolated %s / %s" % (prec, rr["iters"], rr3["iters"], rx["iters"],
                                                                   rx3["iters"]))
    assert rx["converged"].all() and rr["converged"].all() and rx3["converged"].all() and rr3["converged"].all()
    top = rr["lam"][0, 9]
    assert np.abs(rx["lam"][0, :10] - rr["lam"][0, :10]).max() < 1e-8 * top
    assert np.abs(rx3["lam"][0, :10] - rr3["lam"][0, :10]).max() < 1e-8 * top
    # the extrapolated block already holds good approximations of most of
    # the converged eigenvalues (next to the spurious near-gradient ones)
    near = [np.abs(lam_x[0] - l).min() < 1e-3 * top for l in rx["lam"][0, 1:10]]
    assert sum(near) >= 5, (lam_x[0], rx["lam"][0, :10])
