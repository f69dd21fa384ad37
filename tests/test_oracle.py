"""Pins the CPU oracle (oracle/) to the mathematics of the reference's hot path.

Each check is chosen so that a plausible mistake in the oracle fails it: an
independent quadrature assembly (tests/fem_reference.py), closed-form
eigenvalues, the invariants the reference's test programs print (test_edge2d.c,
test_node2d.c, test_jumps2d.c, test_mgedge2d.c), dense eigensolves and finite
differences.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
import fem_reference as F

BLOCH = [(0.0, 0.0), (0.3, -1.1), (np.pi, np.pi), (np.pi, 0.0), (2.0, 0.7)]


def rnd(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def random_eta(nx, ny, seed):
    r = np.random.default_rng(seed)
    return 1.0 / r.uniform(1.0, 12.0, size=(ny, nx))


# ---------------------------------------------------------------- grid2d
def test_setpattern_corner_sampling():
    n = 20
    eta = O.setpattern(n, n, 1.0, circles=[(8.9, 0.5, 0.5, 0.2)])
    for j in range(n):
        for i in range(n):
            x, y = i * (1.0 / n), j * (1.0 / n)
            inside = (x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) <= 0.2 * 0.2
            assert eta[j, i] == (1 / 8.9 if inside else 1.0)
    # rectangles first, circles override (grid2d.c:193-217)
    eta = O.setpattern(n, n, 2.0, rects=[(4.0, 0.0, 0.0, 0.5, 1.0)], circles=[(8.0, 0.25, 0.5, 0.1)])
    assert eta[10, 5] == 1 / 8.0
    assert eta[0, 0] == 1 / 4.0
    assert eta[0, 15] == 1 / 2.0


def test_coarsen_is_mean_of_inverse_eps():
    eta = random_eta(8, 8, 1)
    g = O.Grid2D(eta)
    gc = g.coarse()
    ref = 0.25 * (eta[0::2, 0::2] + eta[0::2, 1::2] + eta[1::2, 0::2] + eta[1::2, 1::2])
    assert np.allclose(gc.eta, ref, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- edge operators
@pytest.mark.parametrize("kx,ky", BLOCH)
def test_dense_matches_quadrature_fem(kx, ky):
    nx, ny = 6, 4
    hx, hy = 0.7 / nx, 0.45 / ny
    eta = random_eta(nx, ny, 2)
    g = O.Grid2D(eta, hx, hy, kx, ky)
    K, M = F.edge_matrices(eta, hx, hy, g.xf, g.yf)
    assert np.allclose(O.dense_edge(g, 1, 0), K, atol=1e-12 * np.abs(K).max())
    assert np.allclose(O.dense_edge(g, 0, 1), M, atol=1e-13)
    A = O.dense_edge(g, 0.7 - 0.2j, 1.3 + 0.4j)
    assert np.allclose(A, (0.7 - 0.2j) * K + (1.3 + 0.4j) * M, atol=1e-12 * np.abs(K).max())


@pytest.mark.parametrize("kx,ky", BLOCH)
def test_addeval_equals_dense(kx, ky):
    nx, ny = 8, 6
    eta = random_eta(nx, ny, 3)
    g = O.Grid2D(eta, 1 / nx, 1.3 / ny, kx, ky)
    x, y0 = rnd(g.nedge, 4), rnd(g.nedge, 5)
    al, be = 1.2 - 0.3j, 0.4 + 0.9j
    y = O.addeval(g, al, be, x, y0)
    ref = y0 + O.dense_edge(g, al, be) @ x
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_addeval_hermitian_and_energyprod():
    n = 8
    g = O.Grid2D(O.rod_eta(n), kx=0.3, ky=-1.1)
    x, y = rnd(g.nedge, 6), rnd(g.nedge, 7)
    ax = O.addeval(g, 1, 1, x)
    ay = O.addeval(g, 1, 1, y)
    # self-adjointness (test_edge2d.c:178-195)
    assert abs(np.vdot(y, ax) - np.vdot(ay, x)) < 1e-12 * abs(np.vdot(y, ax))
    # energyprod = x^H A y (test_edge2d.c:199-209)
    assert abs(O.energyprod(g, 1, 1, x, y) - np.vdot(x, ay)) < 1e-12 * abs(np.vdot(x, ay))
    assert abs(O.massprod(g, x, y) - np.vdot(x, O.addeval(g, 0, 1, y))) < 1e-12


def test_mass_only_reduces_to_tridiagonal():
    """With alpha = 0 addeval is the 4-1 Toeplitz mass stencil per line."""
    nx, ny = 6, 6
    g = O.Grid2D(np.ones((ny, nx)), kx=0.0, ky=0.0)
    x = rnd(g.nedge, 8)
    y = O.addeval(g, 0, 1, x)
    X = x[:36].reshape(ny, nx)
    ref = (g.hy / g.hx / 6) * (4 * X + np.roll(X, 1, 0) + np.roll(X, -1, 0))
    assert np.allclose(y[:36].reshape(ny, nx), ref)
    Y = x[36:].reshape(ny, nx)
    ref = (g.hx / g.hy / 6) * (4 * Y + np.roll(Y, 1, 1) + np.roll(Y, -1, 1))
    assert np.allclose(y[36:].reshape(ny, nx), ref)


@pytest.mark.parametrize("kx,ky", [(0.3, -1.1), (np.pi, 0.5)])
def test_empty_lattice_closed_form(kx, ky):
    nx, ny = 8, 6
    hx, hy = 1 / nx, 0.8 / ny
    g = O.Grid2D(np.full((ny, nx), 0.5), hx, hy, kx, ky)
    w = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1), eigvals_only=True)
    ref = np.sort(np.r_[F.empty_lattice_eigs(nx, ny, hx, hy, kx, ky, 0.5), np.zeros(nx * ny)])
    assert np.abs(w - ref).max() < 1e-9 * ref.max()


def test_reference_eta_quirk():
    """The reference's addeval/energyprod read eps of the cell above for the
    cell below an x-edge in the top row and in the last column.  For the rod
    inputs eta is constant there and the quirk changes nothing bit for bit;
    for a random eta it makes the operator non-Hermitian."""
    for n in (16, 64):
        g = O.Grid2D(O.rod_eta(n), kx=0.3, ky=-1.1)
        x = rnd(g.nedge, 9)
        assert np.array_equal(O.addeval(g, 1, 1, x), O.addeval(g, 1, 1, x, quirk=True))
    g = O.Grid2D(random_eta(8, 8, 10), kx=0.3, ky=-1.1)
    x, y = rnd(g.nedge, 11), rnd(g.nedge, 12)
    lhs = np.vdot(y, O.addeval(g, 1, 1, x, quirk=True))
    rhs = np.vdot(O.addeval(g, 1, 1, y, quirk=True), x)
    assert abs(lhs - rhs) > 1e-6 * abs(lhs)


def test_reference_eta_quirk_shifts_random_field_spectrum():
    """How far the reference's seam-row eta reading moves the spectrum of the
    adjoint workload's kind of field (the thresholded Gaussian random field
    generator, seed 4, sigma scaled to the grid), measured on a 24 x 24 grid
    where the non-Hermitian generalised eigenproblem is solved densely: the
    ten lowest non-zero eigenvalues become complex and their real parts move
    by percents.  The GPU path implements the Hermitian operator (the element
    assembly), so on such fields its eigenvalues are those of the Hermitian
    pencil, not the reference's."""
    import bench
    n = 24
    eta = bench.random_field_eta(n, seed=4, sigma=6.0 * n / 256)
    assert 0.2 < (eta < 1).mean() < 0.8  # both materials, so eta jumps across the seam
    g = O.Grid2D(eta, kx=2.7386, ky=0.0)
    K, M = O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1)
    E = 2 * n * n
    Kq = np.stack([O.addeval(g, 1.0, 0.0, col, quirk=True) for col in np.eye(E)], axis=1)
    assert np.abs(Kq - Kq.conj().T).max() > 0.1 * np.abs(Kq).max()  # not Hermitian
    w = sl.eigh(K, M, eigvals_only=True)
    w = w[w > 1e-6][:10]
    wq = sl.eig(Kq, M, right=False)
    wq = wq[np.abs(wq) > 1e-6]
    wq = wq[np.argsort(wq.real)][:10]
    shift = np.abs(wq.real - w) / w
    imag = np.abs(wq.imag) / np.abs(wq.real)
    print("quirk: real-part shift %s, |imag|/real %s" % (np.round(shift, 4), np.round(imag, 4)))
    assert shift.max() > 1e-3 and imag.max() > 1e-3


@pytest.mark.parametrize("kx,ky", [(0.3, -1.1), (0.0, 0.0), (np.pi, 0.5)])
def test_householder_orthonormalisation(kx, ky):
    """The oracle's Householder M-orthonormalisation (edge2d.c:6154-6296):
    M-orthonormal outputs with nested spans (Q^H M X upper triangular and
    X = Q R), each output equal up to a unit phase to the M-weighted
    Cholesky-QR basis of the inputs, and the closed form for the unit vectors
    themselves (each is mapped to minus itself)."""
    g = O.Grid2D(random_eta(8, 6, 4), 1 / 8, 0.9 / 6, kx, ky)
    Mm = O.dense_edge(g, 0, 1)
    k = 7
    r = np.random.default_rng(21)
    X = r.standard_normal((k, g.nedge)) + 1j * r.standard_normal((k, g.nedge))
    Q, tau, _ = O.orthonormalize(g, X)
    assert np.abs(Q.conj() @ Mm @ Q.T - np.eye(k)).max() < 1e-12
    R = Q.conj() @ Mm @ X.T
    assert np.abs(np.tril(R, -1)).max() < 1e-12 * np.abs(R).max()
    assert np.abs(Q.T @ R - X.T).max() < 1e-12 * np.abs(X).max()
    G = X.conj() @ Mm @ X.T
    Lc = np.linalg.cholesky(G)                      # G = Lc Lc^H, X^T = Qc^T Lc^H
    Qc = np.linalg.solve(Lc.conj(), X)              # rows: Cholesky-QR basis
    ph = np.einsum("ij,jk,ik->i", Q.conj(), Mm, Qc)  # <Q_i, Qc_i>_M
    assert np.abs(np.abs(ph) - 1).max() < 1e-12
    # unit vectors in, minus unit vectors out (alpha = beta = 1, tau = 1/2)
    U = np.zeros((k, g.nedge), np.complex128)
    for i in range(k):
        a, b = O.unit_indices(i, g.nx, g.ny)
        U[i, a + b * g.nx] = np.sqrt(0.25 / (g.hy / g.hx / 6.0))
    Qu, tu, _ = O.orthonormalize(g, U)
    assert np.abs(Qu + U).max() < 1e-14 and np.abs(tu - 0.5).max() < 1e-14


def test_rayleigh_extrapolation_oracle():
    """The oracle's Rayleigh-Ritz extrapolation (exp_blochsurf.c:332-406) by
    what the Rayleigh-Ritz method fixes: with the exact eigenvectors of the
    new Bloch parameter inside the basis it returns them, with their
    eigenvalues; for any basis its vectors are M-orthonormal, lie in the
    basis' span, carry Rayleigh quotients equal to the returned values, and
    those values bound the exact eigenvalues from above (Courant-Fischer);
    and a basis given in another order or mixed by an invertible matrix gives
    the same values (the method sees only the span)."""
    g = O.Grid2D(random_eta(8, 8, 5), kx=0.7, ky=-0.4)
    K, Mm = O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1)
    w, V = sl.eigh(K + Mm, Mm)
    r = np.random.default_rng(5)
    eigs = 4
    # exact eigenvectors 0..3 hidden in a 9-vector basis with random vectors
    B = np.vstack([V[:, :eigs].T, r.standard_normal((5, g.nedge)) + 1j * r.standard_normal((5, g.nedge))])
    B = (r.standard_normal((9, 9)) + 1j * r.standard_normal((9, 9))) @ B
    e, lam = O.rayleigh_extrapolation(g, B, eigs)
    assert np.abs(lam - w[:eigs]).max() < 1e-10 * w[eigs]
    for j in range(eigs):
        res = (K + Mm) @ e[j] - lam[j] * (Mm @ e[j])
        assert np.linalg.norm(res) < 1e-9 * w[eigs] * np.linalg.norm(Mm @ e[j])
    # any basis: M-orthonormal, in the span, Rayleigh quotients, upper bounds
    B2 = r.standard_normal((6, g.nedge)) + 1j * r.standard_normal((6, g.nedge))
    e2, lam2 = O.rayleigh_extrapolation(g, B2, 3)
    assert np.abs(e2.conj() @ Mm @ e2.T - np.eye(3)).max() < 1e-12
    coef, *_ = np.linalg.lstsq(B2.T, e2.T, rcond=None)
    assert np.abs(B2.T @ coef - e2.T).max() < 1e-10 * np.abs(e2).max()
    rq = np.einsum("ij,jk,ik->i", e2.conj(), K + Mm, e2).real
    assert np.abs(rq - lam2).max() < 1e-10 * lam2.max()
    assert (lam2 >= w[:3] - 1e-10 * w[3]).all()
    e3, lam3 = O.rayleigh_extrapolation(g, (r.standard_normal((6, 6)) + 0.5j * np.eye(6)) @ B2[::-1], 3)
    assert np.abs(lam3 - lam2).max() < 1e-10 * lam2.max()


def test_center_removes_constant_fields():
    n = 8
    g = O.Grid2D(O.rod_eta(n))
    x = rnd(g.nedge, 13)
    c = O.center(g, x)
    assert abs(c[:64].sum()) < 1e-12 and abs(c[64:].sum()) < 1e-12
    # the constant fields are K-null and M-orthogonal to the centred vector
    const = np.r_[np.ones(64), np.zeros(64)]
    assert np.abs(O.addeval(g, 1, 0, const)).max() < 1e-10
    assert abs(O.massprod(g, const, c)) < 1e-12


# ---------------------------------------------------------------- transfers
@pytest.mark.parametrize("kx,ky", BLOCH)
def test_edge_restriction_is_adjoint_and_galerkin(kx, ky):
    nf = 8
    eta = random_eta(nf, nf, 14)
    gf = O.Grid2D(eta, kx=kx, ky=ky)
    gc = gf.coarse()
    nc = gc.nedge
    P = np.stack([O.prolong_edge(gf, np.eye(nc)[k]) for k in range(nc)], axis=1)
    R = np.stack([O.restrict_edge(gf, np.eye(gf.nedge)[k]) for k in range(gf.nedge)], axis=1)
    assert np.allclose(R, P.conj().T, atol=1e-15)
    # Galerkin (test_edge2d.c:543-615): P^H K_f P = K_c (coarse eta = mean), P^H M_f P = M_c
    Kf, Mf = O.dense_edge(gf, 1, 0), O.dense_edge(gf, 0, 1)
    Kc, Mc = O.dense_edge(gc, 1, 0), O.dense_edge(gc, 0, 1)
    assert np.allclose(P.conj().T @ Kf @ P, Kc, atol=1e-10 * np.abs(Kc).max())
    assert np.allclose(P.conj().T @ Mf @ P, Mc, atol=1e-13)


@pytest.mark.parametrize("kx,ky", BLOCH)
def test_node_transfers_and_dense(kx, ky):
    nf = 8
    gf = O.Grid2D(np.ones((nf, nf)), kx=kx, ky=ky)
    gc = gf.coarse()
    A = F.node_laplacian(nf, nf, gf.hx, gf.hy, gf.xf, gf.yf)
    assert np.allclose(O.dense_node(gf, fixnull=False), A, atol=1e-13)
    x = rnd(gf.nnode, 15)
    assert np.allclose(O.addeval_node(gf, 1, x), A @ x, atol=1e-12)
    nc = gc.nnode
    P = np.stack([O.prolong_node(gf, np.eye(nc)[k]) for k in range(nc)], axis=1)
    R = np.stack([O.restrict_node(gf, np.eye(gf.nnode)[k]) for k in range(gf.nnode)], axis=1)
    assert np.allclose(R, P.conj().T, atol=1e-15)
    Ac = O.dense_node(gc, fixnull=False)
    assert np.allclose(P.conj().T @ A @ P, Ac, atol=1e-12)


def test_node_fixnull_only_at_gamma():
    g = O.Grid2D(np.ones((4, 4)))
    A0 = O.dense_node(g, fixnull=False)
    A1 = O.dense_node(g, fixnull=True)
    assert np.allclose(A1 - A0, 0.1 / 256)
    g.set_bloch(0.3, 0.0)
    assert np.array_equal(O.dense_node(g, True), O.dense_node(g, False))


@pytest.mark.parametrize("kx,ky", BLOCH)
def test_gradient_adjoint_and_curl_free(kx, ky):
    n = 8
    g = O.Grid2D(O.rod_eta(n), kx=kx, ky=ky)
    G = F.gradient_matrix(n, n, g.xf, g.yf)
    v, e = rnd(g.nnode, 16), rnd(g.nedge, 17)
    assert np.allclose(O.gradient(g, v), G @ v, atol=1e-14)
    assert np.allclose(O.adjgradient(g, e), G.conj().T @ e, atol=1e-14)
    # K G = 0 and G^H M G = node Laplacian scaled: the exact sequence property
    assert np.abs(O.addeval(g, 1, 0, O.gradient(g, v))).max() < 1e-9
    Ln = F.node_laplacian(n, n, g.hx, g.hy, g.xf, g.yf)
    M = O.dense_edge(g, 0, 1)
    assert np.allclose(G.conj().T @ M @ G, Ln, atol=1e-12)


# ---------------------------------------------------------------- smoother
def _patch_dofs(n, i, j, xf, yf):
    """Global indices and phases of the star DOFs of node (i,j)."""
    N = n * n
    out = []
    for (ii, jj, off) in ((i - 1, j, 0), (i, j, 0), (i, j - 1, N), (i, j, N)):
        ph = 1.0 + 0j
        if ii < 0:
            ii += n
            ph *= np.conj(xf)
        if jj < 0:
            jj += n
            ph *= np.conj(yf)
        out.append((off + ii + jj * n, ph))
    return out


@pytest.mark.parametrize("i,j", [(3, 4), (0, 0), (0, 5), (5, 0)])
def test_patch_solve_matrix(i, j):
    """The AFW patch update delta solves (K_patch + 2 M_patch) delta = r_patch
    with r = b - (K + M) x exact (edge2d.c:2029-2183)."""
    n = 8
    g = O.Grid2D(random_eta(n, n, 18), kx=0.3, ky=-1.1)
    b, x = rnd(g.nedge, 19), rnd(g.nedge, 20)
    K, M = O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1)
    x1 = O.patch_solve(g, b, x, i, j)
    dofs = _patch_dofs(n, i, j, g.xf, g.yf)
    idx = [d for d, _ in dofs]
    ph = np.array([p for _, p in dofs])
    changed = np.flatnonzero(x1 != x)
    assert set(changed) <= set(idx)
    delta = ph * (x1 - x)[idx]           # update in the unwrapped (local) frame
    r = (b - (K + M) @ x)[idx] * ph      # local residual in the local frame
    Kp = np.conj(ph)[:, None] * K[np.ix_(idx, idx)] * ph[None, :]
    Mp = np.conj(ph)[:, None] * M[np.ix_(idx, idx)] * ph[None, :]
    # rows in the local frame: ph * (row of the global matrix)
    Kp = ph[:, None] * K[np.ix_(idx, idx)] * np.conj(ph)[None, :]
    Mp = ph[:, None] * M[np.ix_(idx, idx)] * np.conj(ph)[None, :]
    assert np.allclose((Kp + 2 * Mp) @ delta, r, atol=1e-10 * np.abs(r).max())


@pytest.mark.parametrize("order", [0, 1])
def test_gs_forward_backward_adjoint(order):
    """Forward and backward sweeps are adjoint (test_edge2d.c:243-262)."""
    n = 8
    g = O.Grid2D(random_eta(n, n, 21), kx=0.3, ky=-1.1)
    r1, r2 = rnd(g.nedge, 22), rnd(g.nedge, 23)
    z = np.zeros(g.nedge, complex)
    f = O.gs_edge(g, r1, z, backward=False, colour_order=order)
    bk = O.gs_edge(g, r2, z, backward=True, colour_order=order)
    assert abs(np.vdot(r2, f) - np.vdot(bk, r1)) < 1e-11 * abs(np.vdot(r2, f))
    fn = O.gs_node(g, r1[:64], z[:64], backward=False, colour_order=order)
    bn = O.gs_node(g, r2[:64], z[:64], backward=True, colour_order=order)
    assert abs(np.vdot(r2[:64], fn) - np.vdot(bn, r1[:64])) < 1e-11 * abs(np.vdot(r2[:64], fn))


def test_four_colour_sweep_is_order_independent_within_colour():
    """Patches of one colour touch disjoint unknowns and read nothing another
    patch of the colour writes, so the colour phase equals a Jacobi step over
    the colour: applying the patches of one colour in two different orders
    gives the same result bit for bit."""
    n = 8
    g = O.Grid2D(random_eta(n, n, 24), kx=0.3, ky=-1.1)
    b, x = rnd(g.nedge, 25), rnd(g.nedge, 26)
    nodes = [(i, j) for j in range(0, n, 2) for i in range(1, n, 2)]
    xa = x.copy()
    for (i, j) in nodes:
        xa = O.patch_solve(g, b, xa, i, j)
    xb = x.copy()
    for (i, j) in reversed(nodes):
        xb = O.patch_solve(g, b, xb, i, j)
    assert np.array_equal(xa, xb)


# ---------------------------------------------------------------- multigrid
@pytest.mark.parametrize("order", [0, 1])
def test_vcycle_hermitian_and_contracting(order):
    """V-cycle symmetry (test_jumps2d.c:146-184) and rate (test_mgedge2d.c)."""
    mg = O.Multigrid(4, 2, O.rod_eta(16), colour_order=order)
    mg.set_bloch(0.3, -1.1)
    g = mg.grid()
    r1, r2 = rnd(g.nedge, 27), rnd(g.nedge, 28)
    v1, v2 = mg.vcycle_edge(r1), mg.vcycle_edge(r2)
    assert abs(np.vdot(r2, v1) - np.vdot(v2, r1)) < 1e-11 * abs(np.vdot(r2, v1))
    n1, n2 = mg.vcycle_node(r1[:256]), mg.vcycle_node(r2[:256])
    assert abs(np.vdot(r2[:256], n1) - np.vdot(n2, r1[:256])) < 1e-11 * abs(np.vdot(r2[:256], n1))
    x = np.zeros(g.nedge, complex)
    res = []
    for _ in range(6):
        x = mg.vcycle_edge(r1, x)
        res.append(np.linalg.norm(r1 - O.addeval(g, 1, 1, x)))
    assert res[-1] / res[-2] < 0.4
    xn = np.zeros(256, complex)
    res = []
    for _ in range(6):
        xn = mg.vcycle_node(r1[:256], xn)
        res.append(np.linalg.norm(r1[:256] - O.addeval_node(g, 1, xn)))
    assert res[-1] / res[-2] < 0.2


def test_gradient_lifting_eliminates_divergence():
    """After lifting, G^H M e is much smaller (test_node2d.c:494-556)."""
    mg = O.Multigrid(4, 3, O.rod_eta(32))
    mg.set_bloch(0.3, -1.1)
    g = mg.grid()
    e = rnd(g.nedge, 29)
    div0 = np.linalg.norm(O.adjgradient(g, O.addeval(g, 0, 1, e)))
    e1 = mg.lift(e)
    div1 = np.linalg.norm(O.adjgradient(g, O.addeval(g, 0, 1, e1)))
    assert div1 < 1e-4 * div0


# ---------------------------------------------------------------- dense linear algebra
def test_potrf_potrs_hegv_against_scipy():
    n = 12
    r = np.random.default_rng(30)
    X = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    B = X @ X.conj().T + n * np.eye(n)
    Y = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    A = Y + Y.conj().T
    L, info = O.potrf(B)
    assert info == 0
    assert np.allclose(L, sl.cholesky(B, lower=True))
    b = rnd(n, 31)
    assert np.allclose(O.potrs(L, b), np.linalg.solve(B, b))
    w, V = O.hegv(np.tril(A), np.tril(B))
    assert np.allclose(w, sl.eigh(A, B, eigvals_only=True), rtol=1e-12, atol=1e-12)
    assert np.allclose(V.conj().T @ B @ V, np.eye(n), atol=1e-12)
    assert np.allclose(A @ V, B @ V * w, atol=1e-10)
    # non-definite B is reported, as zhegv does
    with pytest.raises(np.linalg.LinAlgError):
        O.hegv(np.tril(A), np.tril(A))


# ---------------------------------------------------------------- eigensolver
@pytest.mark.parametrize("kx,ky", [(0.3, -1.1), (0.0, 0.0), (np.pi, np.pi)])
def test_block_pinvit_matches_dense_eigs(kx, ky):
    mg = O.Multigrid(4, 2, O.rod_eta(16))
    mg.set_bloch(kx, ky)
    g = mg.grid()
    lam, E, rel, ab, it = mg.block_solve(O.random_block(12, g.nedge, 32), 10, tol=1e-9, maxit=300)
    assert rel[:10].max() <= 1e-9
    w = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1), eigvals_only=True)
    w = w[np.abs(w) > 1e-6][:10]
    assert np.abs(lam[:10] - w).max() / w.max() < 1e-10
    # Ritz vectors are M-orthonormal
    M = O.dense_edge(g, 0, 1)
    G = E[:10].conj() @ M @ E[:10].T
    assert np.allclose(G, np.eye(10), atol=1e-8)


def test_block_pinvit_empty_lattice_and_reference_order():
    mg = O.Multigrid(4, 2, np.ones((16, 16)), colour_order=0)
    mg.set_bloch(0.3, -1.1)
    g = mg.grid()
    lam, E, rel, ab, it = mg.block_solve(O.random_block(12, g.nedge, 33), 10, tol=1e-9, maxit=300)
    ref = F.empty_lattice_eigs(16, 16, 1 / 16, 1 / 16, 0.3, -1.1)[:10]
    assert np.abs(lam[:10] - ref).max() / ref.max() < 1e-10


def test_eta_gradient_matches_finite_differences():
    """d lambda_i / d eta_c against central differences of dense eigenvalues."""
    n = 8
    eta = O.rod_eta(n)
    g = O.Grid2D(eta, kx=0.3, ky=-1.1)
    w, V = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1))
    nz = np.flatnonzero(np.abs(w) > 1e-6)
    i = nz[3]
    grad = O.eta_gradient(g, V[:, i])
    h = 1e-5
    for (cj, ci) in [(0, 0), (3, 4), (7, 2)]:
        ep, em = eta.copy(), eta.copy()
        ep[cj, ci] += h
        em[cj, ci] -= h
        gp, gm = O.Grid2D(ep, kx=0.3, ky=-1.1), O.Grid2D(em, kx=0.3, ky=-1.1)
        wp = sl.eigh(O.dense_edge(gp, 1, 0), O.dense_edge(gp, 0, 1), eigvals_only=True)[i]
        wm = sl.eigh(O.dense_edge(gm, 1, 0), O.dense_edge(gm, 0, 1), eigvals_only=True)[i]
        fd = (wp - wm) / (2 * h)
        assert abs(fd - grad[cj, ci]) < 1e-6 * max(1.0, abs(fd))


def test_oracle_quadrature_identities():
    """What fixes the oracle's restatements of the quadrature routines:
    (edge) for a constant field at Gamma the L2 functional equals M times the
    interpolant (the rule is exact on constants and the edge mass stencil
    integrates the interpolant exactly), the interpolant's L2 error is zero and
    its L2 norm is |c|; (node) the 2 x 2 Gauss rule integrates products of
    bilinears exactly, so the node L2 norm squared is v^H M_Q1 v with the Q1
    mass matrix of the Bloch grid, and the functional of 1 at Gamma is hx hy per
    node; (both) smooth fields are approximated to second order in h."""
    nx = ny = 8
    h = 1.0 / nx
    c = (0.7 - 0.2j, -0.3 + 1.1j)
    const = lambda p: c
    g = O.Grid2D(np.ones((ny, nx)))
    Mm = O.dense_edge(g, 0.0, 1.0)
    vi = O.interpolate_edge(const, nx, ny, h, h)
    assert np.abs(O.l2functional_edge(const, nx, ny, h, h, 1.0, 1.0) - Mm @ vi).max() < 1e-15
    assert O.l2norm_edge(vi, const, nx, ny, h, h, 1.0, 1.0) < 1e-14
    assert abs(O.l2norm_edge(vi, None, nx, ny, h, h, 1.0, 1.0) - np.sqrt(abs(c[0]) ** 2 + abs(c[1]) ** 2)) < 1e-14
    # node: the Q1 mass matrix with Bloch factors, from its element matrix
    for xf, yf in ((1.0, 1.0), (np.exp(0.4j), np.exp(-1.3j))):
        rng = np.random.default_rng(1)
        v = rng.standard_normal(nx * ny) + 1j * rng.standard_normal(nx * ny)
        Me = np.array([[4, 2, 2, 1], [2, 4, 1, 2], [2, 1, 4, 2], [1, 2, 2, 4]]) * h * h / 36.0
        V = v.reshape(ny, nx)
        q = 0.0
        for j in range(ny):
            for i in range(nx):
                u = []
                for b in (0, 1):
                    for a in (0, 1):
                        f = (xf if i + a == nx else 1.0) * (yf if j + b == ny else 1.0)
                        u.append(f * V[(j + b) % ny, (i + a) % nx])
                u = np.array(u)  # corners (0,0), (1,0), (0,1), (1,1)
                q += np.vdot(u, Me @ u).real
        assert abs(O.l2norm_node(v, None, nx, ny, h, h, xf, yf) ** 2 - q) < 1e-13 * q
    assert np.abs(O.l2functional_node(lambda p: 1.0, nx, ny, h, h, 1.0, 1.0) - h * h).max() < 1e-16
    # second order: interpolation errors of smooth periodic fields
    fe = lambda p: (np.cos(2 * np.pi * p[1]), np.sin(2 * np.pi * p[0]))
    fn = lambda p: np.exp(2j * np.pi * (p[0] + 2 * p[1]))
    err = []
    for n in (16, 32):
        hh = 1.0 / n
        err.append((O.l2norm_edge(O.interpolate_edge(fe, n, n, hh, hh), fe, n, n, hh, hh, 1.0, 1.0),
                    O.l2norm_node(np.array([fn((i * hh, j * hh)) for j in range(n) for i in range(n)]), fn,
                                  n, n, hh, hh, 1.0, 1.0)))
    assert err[0][0] / err[1][0] > 3.5 and err[0][1] / err[1][1] > 3.5, err


def test_oracle_simple_smoother_is_gauss_seidel():
    """The dense pointwise Gauss-Seidel, in the four-colour order: the
    solution of (K + M) x = b is its fixed point and every sweep lowers the
    energy norm of the error (A is Hermitian positive definite); colours are
    independent sets, so the order within a colour does not matter (reversed
    order, same result)."""
    n = 8
    g = O.Grid2D(O.rod_eta(n), kx=0.3, ky=-1.1)
    A = O.dense_edge(g, 1.0, 1.0)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(2 * n * n) + 1j * rng.standard_normal(2 * n * n)
    order = O.simple_colour_order(n, n)
    N = n * n
    cols = [order[k * N // 2:(k + 1) * N // 2] for k in range(4)]
    x0 = np.zeros(2 * n * n, complex)
    a1 = O.gs_simple_dense(A, b, x0, order)
    a2 = O.gs_simple_dense(A, b, x0, [e for c in cols for e in c[::-1]])
    assert np.abs(a1 - a2).max() < 1e-13 * np.abs(a1).max()
    xs = np.linalg.solve(A, b)
    assert np.abs(O.gs_simple_dense(A, b, xs, order) - xs).max() < 1e-12 * np.abs(xs).max()
    x, errs = x0, []
    for _ in range(10):
        x = O.gs_simple_dense(A, b, x, order)
        d = x - xs
        errs.append(np.vdot(d, A @ d).real)
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:])), errs
