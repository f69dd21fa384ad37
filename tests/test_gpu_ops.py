"""GPU parity of the reference module API (gfx950 kernels via the C-ABI)
against the CPU oracle on the same seeded inputs."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BLOCH = [(0.0, 0.0), (0.3, -1.1), (np.pi, np.pi), (2.0, 0.7)]


@pytest.fixture(scope="module")
def M():
    import maxwell_amd
    return maxwell_amd


def rnd(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def random_eta(n, seed):
    return 1.0 / np.random.default_rng(seed).uniform(1.0, 12.0, size=(n, n))


def close(a, b, rtol):
    return np.abs(a - b).max() <= rtol * max(np.abs(b).max(), 1e-300)


def pair(M, eta, kx, ky):
    g = M.Grid(eta, kx, ky)
    o = O.Grid2D(eta, kx=kx, ky=ky)
    o.set_factors(*g.factors)
    return g, o


@pytest.mark.parametrize("kx,ky", BLOCH)
def test_addeval_energy_mass_dot(M, kx, ky):
    n = 16
    g, o = pair(M, random_eta(n, 1), kx, ky)
    x, y0 = rnd(o.nedge, 2), rnd(o.nedge, 3)
    X, Y = M.Edge(g, x), M.Edge(g, y0)
    al, be = 1.3 - 0.2j, 0.7 + 0.4j
    M.lib.addeval_edge2d(M.Cplx.of(al), M.Cplx.of(be), X.h, Y.h)
    assert close(Y.get(), O.addeval(o, al, be, x, y0), 1e-13)
    Y.set(y0)
    e = M.lib.energyprod_edge2d(M.Cplx.of(al), M.Cplx.of(be), X.h, Y.h).value()
    assert abs(e - O.energyprod(o, al, be, x, y0)) < 1e-12 * abs(e)
    m = M.lib.massprod_edge2d(X.h, Y.h).value()
    assert abs(m - O.massprod(o, x, y0)) < 1e-12 * abs(m)
    d = M.lib.dotprod_edge2d(X.h, Y.h).value()
    assert abs(d - np.vdot(x, y0)) < 1e-12 * abs(d)
    assert abs(M.lib.norm2_edge2d(X.h) - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)
    assert abs(M.lib.normmax_edge2d(X.h) - np.abs(x).max()) < 1e-14 * np.abs(x).max()


@pytest.mark.parametrize("kx,ky", BLOCH)
def test_afw_sweeps_match_oracle_four_colour(M, kx, ky):
    n = 16
    g, o = pair(M, random_eta(n, 4), kx, ky)
    b, x = rnd(o.nedge, 5), rnd(o.nedge, 6)
    B, X = M.Edge(g, b), M.Edge(g, x)
    one = M.Cplx.of(1.0)
    M.lib.gsforward_edge2d(one, one, B.h, X.h)
    ref = O.gs_edge(o, b, x, backward=False, colour_order=1)
    assert close(X.get(), ref, 1e-12)
    M.lib.gsbackward_edge2d(one, one, B.h, X.h)
    ref = O.gs_edge(o, b, ref, backward=True, colour_order=1)
    assert close(X.get(), ref, 1e-12)


@pytest.mark.parametrize("kx,ky", BLOCH)
def test_transfers(M, kx, ky):
    n = 16
    g, o = pair(M, random_eta(n, 7), kx, ky)
    gc = g.coarse()
    gc.set_bloch(kx, ky)
    c, f = rnd(o.nedge // 4, 8), rnd(o.nedge, 9)
    Cv, F = M.Edge(gc, c), M.Edge(g, f)
    M.lib.prolongation_edge2d(M.Cplx.of(0.5 - 0.25j), Cv.h, F.h)
    assert close(F.get(), O.prolong_edge(o, c, f, alpha=0.5 - 0.25j), 1e-14)
    F.set(f)
    M.lib.restriction_edge2d(M.Cplx.of(1.5), F.h, Cv.h)
    assert close(Cv.get(), O.restrict_edge(o, f, c, alpha=1.5), 1e-14)
    # node transfers
    cn, fn = rnd(o.nnode // 4, 10), rnd(o.nnode, 11)
    Cn, Fn = M.Node(gc, cn), M.Node(g, fn)
    M.lib.prolongation_node2d(M.Cplx.of(2.0), Cn.h, Fn.h)
    assert close(Fn.get(), O.prolong_node(o, cn, fn, alpha=2.0), 1e-14)
    Fn.set(fn)
    M.lib.restriction_node2d(M.Cplx.of(1.0), Fn.h, Cn.h)
    assert close(Cn.get(), O.restrict_node(o, fn, cn), 1e-14)


@pytest.mark.parametrize("kx,ky", BLOCH)
def test_node_ops_and_gradients(M, kx, ky):
    n = 16
    g, o = pair(M, random_eta(n, 12), kx, ky)
    v, y0, e = rnd(o.nnode, 13), rnd(o.nnode, 14), rnd(o.nedge, 15)
    Vn, Yn, E = M.Node(g, v), M.Node(g, y0), M.Edge(g, e)
    M.lib.addeval_node2d(M.Cplx.of(0.3 + 0.1j), Vn.h, Yn.h)
    assert close(Yn.get(), O.addeval_node(o, 0.3 + 0.1j, v, y0), 1e-13)
    Yn.set(y0)
    M.lib.gsforward_node2d(M.Cplx.of(1.0), Yn.h, Vn.h)
    ref = O.gs_node(o, y0, v, backward=False, colour_order=1)
    assert close(Vn.get(), ref, 1e-13)
    Vn.set(v)
    M.lib.gradient_node2d(M.Cplx.of(-1.0), Vn.h, E.h)
    assert close(E.get(), O.gradient(o, v, e, alpha=-1.0), 1e-14)
    E.set(e)
    M.lib.adjgradient_node2d(M.Cplx.of(1.0), E.h, Yn.h)
    assert close(Yn.get(), O.adjgradient(o, e, y0), 1e-14)


def hierarchy(M, eta, n0, L, kx, ky):
    grids = [None] * (L + 1)
    grids[L] = M.Grid(eta, kx, ky)
    for l in range(L - 1, -1, -1):
        grids[l] = grids[l + 1].coarse()
        grids[l].set_bloch(kx, ky)
    ne, nn = 2 * n0 * n0, n0 * n0
    Ae, An = M.lib.new_matrix(ne, ne), M.lib.new_matrix(nn, nn)
    one = M.Cplx.of(1.0)
    M.lib.densematrix_edge2d(one, one, grids[0].g, Ae)
    assert M.lib.potrf_matrix(Ae) == 0
    M.lib.densematrix_node2d(True, grids[0].g, An)
    assert M.lib.potrf_matrix(An) == 0
    return grids, Ae, An


def arr(objs):
    return (C.c_void_p * len(objs))(*[o.h for o in objs])


@pytest.mark.parametrize("kx,ky", [(0.3, -1.1), (0.0, 0.0)])
def test_vcycles_match_oracle(M, kx, ky):
    n0, L = 4, 3
    eta = O.rod_eta(n0 << L)
    grids, Ae, An = hierarchy(M, eta, n0, L, kx, ky)
    mg = O.Multigrid(n0, L, eta, colour_order=1)
    mg.set_bloch(kx, ky)
    o = mg.grid()
    b = rnd(o.nedge, 16)
    B = [M.Edge(gr) for gr in grids]
    X = [M.Edge(gr) for gr in grids]
    D = [M.Edge(gr) for gr in grids]
    B[L].set(b)
    X[L].set(np.zeros(o.nedge, complex))
    M.lib.vcycle_edge2d(L, 2, Ae, arr(B), arr(X), arr(D))
    M.lib.vcycle_edge2d(L, 2, Ae, arr(B), arr(X), arr(D))
    ref = mg.vcycle_edge(b, mg.vcycle_edge(b))
    assert close(X[L].get(), ref, 1e-11)
    bn = rnd(o.nnode, 17)
    Bn = [M.Node(gr) for gr in grids]
    Xn = [M.Node(gr) for gr in grids]
    Dn = [M.Node(gr) for gr in grids]
    Bn[L].set(bn)
    Xn[L].set(np.zeros(o.nnode, complex))
    M.lib.vcycle_node2d(L, 2, An, arr(Bn), arr(Xn), arr(Dn))
    assert close(Xn[L].get(), mg.vcycle_node(bn), 1e-11)
    # PINVIT vector step (edge2d.c:6298-6328)
    e = rnd(o.nedge, 18)
    E = M.Edge(grids[L], e)
    lam = 30.0
    M.lib.pinvit_edge2d(L, 2, 2, 5, lam, E.h, Ae, arr(B), arr(X), arr(D), An, arr(Bn), arr(Xn), arr(Dn))
    ref = mg.pinvit_vector(lam, e) if (kx, ky) != (0.0, 0.0) else None
    if ref is not None:
        assert close(E.get(), ref, 1e-10)
    M.lib.del_matrix(Ae)
    M.lib.del_matrix(An)


def test_center_and_nullprod(M):
    n = 16
    g, o = pair(M, O.rod_eta(n), 0.0, 0.0)
    x = rnd(o.nedge, 19)
    X = M.Edge(g, x)
    a, b = M.Cplx(), M.Cplx()
    M.lib.nullprod_edge2d(X.h, C.byref(a), C.byref(b))
    N = o.nnode
    assert abs(a.value() - x[:N].sum() / n) < 1e-12 * abs(a.value())
    assert abs(b.value() - x[N:].sum() / n) < 1e-12 * abs(b.value())
    M.lib.center_edge2d(X.h)
    assert close(X.get(), O.center(o, x), 1e-13)
