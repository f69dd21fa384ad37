"""The LOBPCG form of the batched solver (mx_config.method = MX_METHOD_LOBPCG):
Rayleigh-Ritz over each problem's block E, its preconditioned update W (the
same V-cycles and gradient lifting as PINVIT) and the previous step's
direction P -- the subspace of the reference's LOBPCG program
(test_lobpcg2d.c:247-312, basis [u, e, e_old], whose span equals
[e, w, p]) -- against dense generalised eigensolves of the oracle's
assembly, against the PINVIT form (same eigenvalues, far fewer block steps),
and for the vectors it leaves behind."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

pytestmark = pytest.mark.gpu

KS = [(0.3, -1.1), (0.0, 0.0), (np.pi, np.pi), (np.pi, 0.0), (1.0, 0.5), (0.2, 0.0)]


@pytest.fixture(scope="module")
def M():
    import maxwell_amd
    return maxwell_amd


def dense_eigs(eta, kx, ky, count):
    g = O.Grid2D(eta, kx=kx, ky=ky)
    w = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1), eigvals_only=True)
    return w[np.abs(w) > 1e-6][:count]


# block 11: an odd block takes the general block update (mxh_combine_x)
@pytest.mark.parametrize("prec,block", [(0, 12), (1, 12), (1, 11)])
def test_lobpcg_eigenvalues_match_dense(M, prec, block):
    n0, L, n = 4, 3, 32
    eta = O.rod_eta(n)
    kx = np.array([k[0] for k in KS])
    ky = np.array([k[1] for k in KS])
    S = M.Solver(eta, n0=n0, levels=L, precision=prec, block=block, max_problems=len(KS),
                 method=M.MX_METHOD_LOBPCG)
    r = S.solve(kx, ky, neig=10, tol=1e-9, maxit=300, seed=7)
    S.close()
    print("prec %d steps %s" % (prec, r["iters"].tolist()))
    assert r["converged"].all(), r["iters"]
    assert (r["relres"][:, :10] <= 1e-9).all()
    for p, (a, b) in enumerate(KS):
        ref = dense_eigs(eta, a, b, 10)
        # the dense solve resolves eigenvalues to about eps |K| absolute, so
        # the lowest band next to Gamma (0.02 at k = (0.2, 0)) is compared on
        # the scale of the ten
        err = np.abs(r["lam"][p, :10] - ref) / ref.max()
        assert err.max() < 1e-11, (p, err)


@pytest.mark.parametrize("prec", [0, 1])
def test_lobpcg_needs_fewer_steps_than_pinvit(M, prec):
    """Same preconditioner, same tolerance, same start blocks: LOBPCG reaches
    PINVIT's eigenvalues in well under half the block steps (a CPU model of
    both with the oracle's preconditioner: 15-17 against 59-70 steps)."""
    n0, L, n = 4, 5, 128
    eta = O.rod_eta(n)
    kx, ky = [0.0, 0.7, np.pi, 0.17], [0.0, -0.4, np.pi, 0.0]
    out = {}
    for method in (M.MX_METHOD_PINVIT, M.MX_METHOD_LOBPCG):
        S = M.Solver(eta, n0=n0, levels=L, precision=prec, block=12, max_problems=4, method=method)
        out[method] = S.solve(kx, ky, neig=10, tol=1e-8, maxit=400, seed=2)
        S.close()
    a, b = out[M.MX_METHOD_PINVIT], out[M.MX_METHOD_LOBPCG]
    print("steps pinvit %s lobpcg %s" % (a["iters"].tolist(), b["iters"].tolist()))
    assert a["converged"].all() and b["converged"].all()
    assert np.abs(a["lam"][:, :10] - b["lam"][:, :10]).max() / a["lam"][:, 9].max() < 1e-10
    assert (b["iters"] < 0.5 * a["iters"]).all(), (a["iters"], b["iters"])


@pytest.mark.parametrize("prec", [0, 1])
def test_lobpcg_vectors_are_ritz_vectors(M, prec):
    """Problems finish on steps of both parities; every problem's vectors
    afterwards are its final Ritz vectors (M-orthonormal, with the residuals
    mx_solve reported) and the eigenvectors of the dense pencil.  The Ritz
    basis [E W P] is kept to a scaled-Gram condition of about 3m / 1e-7, so
    the block is M-orthonormal to about 1e-16 times that."""
    n = 32
    eta = O.rod_eta(n)
    ks = [(0.0, 0.0), (0.3, -1.1), (np.pi, np.pi), (1.0, 0.5), (np.pi, 0.0), (0.4, 2.0), (2.2, -0.7),
          (-1.3, 0.9)]
    S = M.Solver(eta, n0=4, levels=3, precision=prec, block=12, max_problems=len(ks), method=M.MX_METHOD_LOBPCG)
    r = S.solve([k[0] for k in ks], [k[1] for k in ks], neig=10, tol=1e-9, maxit=300, seed=8)
    assert r["converged"].all()
    print("steps %s" % r["iters"].tolist())
    for p, (kx, ky) in enumerate(ks):
        g = O.Grid2D(eta, kx=kx, ky=ky)
        K, Mm = O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1)
        E = S.vectors(p)
        G = E.conj() @ Mm @ E.T
        assert np.abs(G - np.eye(12)).max() < 1e-8, (p, r["iters"][p])
        for i in range(10):
            e = E[i]
            me = Mm @ e
            rel = np.linalg.norm(K @ e - r["lam"][p, i] * me) / (abs(r["lam"][p, i]) * np.linalg.norm(me))
            assert rel <= 1e-9 and abs(rel - r["relres"][p, i]) <= 1e-3 * r["relres"][p, i] + 1e-13, \
                (p, i, rel, r["relres"][p, i])
    S.close()


def test_lobpcg_configuration_limits(M):
    """Blocks of at most 16 (the Gram kernel takes 48 vectors per problem),
    locking on (without it the converged pairs of the lowest band next to
    Gamma stalled above the tolerance on the 32^2 rod at k = (0.2, 0): their
    update directions are rounding noise that the Ritz step mixes back)."""
    eta = O.rod_eta(16)
    with pytest.raises(RuntimeError, match="error 1"):
        M.Solver(eta, n0=4, levels=2, block=17, method=M.MX_METHOD_LOBPCG)
    with pytest.raises(RuntimeError, match="error 1"):
        M.Solver(eta, n0=4, levels=2, block=8, lock=False, method=M.MX_METHOD_LOBPCG)
    with pytest.raises(RuntimeError, match="error 1"):
        M.Solver(eta, n0=4, levels=2, block=8, method=5)
    S = M.Solver(eta, n0=4, levels=2, block=8, max_problems=1, method=M.MX_METHOD_LOBPCG)
    E0 = O.random_block(8, 2 * 16 * 16, 3)
    E0[5] = E0[2]
    with pytest.raises(RuntimeError, match="error 4"):
        S.solve([0.3], [-1.1], neig=6, E0=E0)
    r = S.solve([0.3], [-1.1], neig=6, tol=1e-9, maxit=300, seed=2)
    assert r["converged"].all()
    S.close()
