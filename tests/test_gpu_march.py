"""Kernel-level parity of the row-marching smoothers (mxh_afw_march,
mxh_node_march), one sweep or two sweeps per pass, with the oracle's
four-colour sweeps, transfers and residuals, called through the device C-ABI
(include/mxh.h) on grids that span several marching tiles in both directions
and wrap with complex Bloch factors."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_abi import NONE, Bloch, Dev, Level, Vecs, bind, vecs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import maxwell_amd as M
    return bind(M.lib)


def cplx(prec):
    return np.complex128 if prec == 0 else np.complex64


def tolerance(prec, h):
    """A patch matrix of the curl-curl + mass operator has condition ~ 8.9/h^2
    (its local gradient direction sees only the mass), so two exact local
    solves written differently agree to ~eps/h^2 relative; an indexing or
    missing-term error shows up at 1e-6 or above even at FP64."""
    return (2e-15 if prec == 0 else 1e-6) / (h * h)


def close(a, b, prec, h):
    err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
    print("    rel err %.2e (tol %.2e)" % (err, tolerance(prec, h)))
    return err <= tolerance(prec, h)


# (grid n, mesh width h, owned rows per marching tile, kx, ky): 256 spans 3 x 2
# and 3 x 7 tiles; h = 1/16 there keeps FP32 well conditioned
CASES = [(256, 1 / 16, 128, 0.7, -1.3), (256, 1 / 16, 40, 0.0, 0.0), (64, 1 / 64, 128, np.pi, 0.4),
         (16, 1 / 16, 128, -2.0, 1.1), (8, 1 / 8, 6, 0.3, 0.9)]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,h,rows,kx,ky", CASES)
@pytest.mark.parametrize("mode", [0, 1, 2, 5, 8, 10])
@pytest.mark.parametrize("nsweep", [1, 2])
def test_afw_march_matches_oracle(lib, prec, n, h, rows, kx, ky, mode, nsweep):
    print("afw_march prec %d n %d rows %d mode %d nsweep %d" % (prec, n, rows, mode, nsweep), flush=True)
    rng = np.random.default_rng(n + mode)
    eta = O.rod_eta(n)
    g = O.Grid2D(eta, hx=h, hy=h, kx=kx, ky=ky)
    E, Ec = 2 * n * n, n * n // 2
    nv = 2
    b = rng.standard_normal((nv, E)) + 1j * rng.standard_normal((nv, E))
    x = rng.standard_normal((nv, E)) + 1j * rng.standard_normal((nv, E))
    xc = rng.standard_normal((nv, Ec)) + 1j * rng.standard_normal((nv, Ec))
    dt = cplx(prec)
    rt = np.float64 if prec == 0 else np.float32
    db, dx, dxo = Dev(lib, b.astype(dt)), Dev(lib, x.astype(dt)), Dev(lib, np.zeros((nv, E), dt))
    dxc, dbc = Dev(lib, xc.astype(dt)), Dev(lib, np.zeros((nv, Ec), dt))
    deta = Dev(lib, eta.astype(rt))
    lv = Level(n, n, h, h)
    bl = Bloch(np.cos(kx), np.sin(kx), np.cos(ky), np.sin(ky), None, 1, None)
    none = NONE
    xc_in = vecs(dxc, Ec) if mode & 4 else none
    bc = vecs(dbc, Ec) if mode & 8 else none
    xcv = vecs(dxc, Ec) if mode & 8 else none
    err = lib.mxh_afw_march(prec, C.byref(lv), C.byref(bl), nv, vecs(db, E), vecs(dx, E), vecs(dxo, E), deta.p,
                            mode, xc_in, bc, xcv, rows, nsweep)
    if nsweep == 2 and n < 10:
        assert err != 0  # halo of two sweeps (10) larger than the grid
        return
    assert err == 0
    got = dxo.get()
    for v in range(nv):
        x0 = np.zeros(E, complex) if mode & 2 else x[v].astype(dt).astype(complex)
        if mode & 4:
            x0 = O.prolong_edge(g, xc[v].astype(dt).astype(complex), x0)
        ref = x0
        for _ in range(nsweep):
            ref = O.gs_edge(g, b[v].astype(dt).astype(complex), ref, backward=bool(mode & 1), colour_order=1)
        assert close(got[v], ref, prec, h), v
        if mode & 8:
            # residual of the kernel's own sweep result, restricted; FP32 is
            # judged against the size of the operator terms it cancels
            xs = ref if prec == 0 else got[v].astype(complex)
            bb = b[v].astype(dt).astype(complex)
            rc = O.restrict_edge(g, O.addeval(g, -1.0, -1.0, xs, bb))
            scale = np.abs(rc).max() if prec == 0 else np.abs(O.restrict_edge(g, O.addeval(g, 1.0, 1.0, xs))).max()
            err = np.abs(dbc.get()[v] - rc).max() / scale
            print("    restricted residual rel err %.2e" % err)
            assert err <= tolerance(prec, h)
            assert np.abs(dxc.get()[v]).max() == 0


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,h,rows,kx,ky", CASES)
@pytest.mark.parametrize("mode", [0, 1, 2, 5, 8, 10])
@pytest.mark.parametrize("nsweep", [1, 2])
def test_node_march_matches_oracle(lib, prec, n, h, rows, kx, ky, mode, nsweep):
    print("node_march prec %d n %d rows %d mode %d nsweep %d" % (prec, n, rows, mode, nsweep), flush=True)
    rng = np.random.default_rng(7 * n + mode)
    g = O.Grid2D(np.ones((n, n)), hx=h, hy=h, kx=kx, ky=ky)
    N, Nc = n * n, n * n // 4
    nv = 2
    b = rng.standard_normal((nv, N)) + 1j * rng.standard_normal((nv, N))
    x = rng.standard_normal((nv, N)) + 1j * rng.standard_normal((nv, N))
    xc = rng.standard_normal((nv, Nc)) + 1j * rng.standard_normal((nv, Nc))
    dt = cplx(prec)
    db, dx, dxo = Dev(lib, b.astype(dt)), Dev(lib, x.astype(dt)), Dev(lib, np.zeros((nv, N), dt))
    dxc, dbc = Dev(lib, xc.astype(dt)), Dev(lib, np.zeros((nv, Nc), dt))
    lv = Level(n, n, h, h)
    bl = Bloch(np.cos(kx), np.sin(kx), np.cos(ky), np.sin(ky), None, 1, None)
    none = NONE
    xc_in = vecs(dxc, Nc) if mode & 4 else none
    bc = vecs(dbc, Nc) if mode & 8 else none
    xcv = vecs(dxc, Nc) if mode & 8 else none
    err = lib.mxh_node_march(prec, C.byref(lv), C.byref(bl), nv, vecs(db, N), vecs(dx, N), vecs(dxo, N), mode,
                             xc_in, bc, xcv, rows, nsweep)
    if nsweep == 2 and n < 10:
        assert err != 0
        return
    assert err == 0
    got = dxo.get()
    for v in range(nv):
        x0 = np.zeros(N, complex) if mode & 2 else x[v].astype(dt).astype(complex)
        if mode & 4:
            x0 = O.prolong_node(g, xc[v].astype(dt).astype(complex), x0)
        ref = x0
        for _ in range(nsweep):
            ref = O.gs_node(g, b[v].astype(dt).astype(complex), ref, backward=bool(mode & 1), colour_order=1)
        assert close(got[v], ref, prec, h), v
        if mode & 8:
            xs = ref if prec == 0 else got[v].astype(complex)
            rc = O.restrict_node(g, b[v].astype(dt).astype(complex) - O.addeval_node(g, 1.0, xs))
            scale = np.abs(rc).max() if prec == 0 else np.abs(O.restrict_node(g, O.addeval_node(g, 1.0, xs))).max()
            err = np.abs(dbc.get()[v] - rc).max() / scale
            print("    restricted residual rel err %.2e" % err)
            assert err <= tolerance(prec, h)
            assert np.abs(dxc.get()[v]).max() == 0


def test_march_rejects_bad_arguments(lib):
    lv = Level(16, 16, 1 / 16, 1 / 16)
    bl = Bloch(1, 0, 1, 0, None, 1, None)
    d = Dev(lib, np.zeros(2 * 256, np.complex128))
    v = vecs(d, 512)
    none = NONE
    # backward with residual, x aliasing the output, missing coarse vectors
    assert lib.mxh_afw_march(0, C.byref(lv), C.byref(bl), 1, v, v, v, d.p, 0, none, none, none, 0, 1) != 0
    assert lib.mxh_node_march(0, C.byref(lv), C.byref(bl), 1, v, v, Vecs(d.p, 512, None), 9, none, none, none,
                              0, 1) != 0
    lv_odd = Level(15, 16, 1 / 15, 1 / 16)
    assert lib.mxh_node_march(0, C.byref(lv_odd), C.byref(bl), 1, v, Vecs(None, 512, None), v, 0, none, none, none,
                              0, 1) != 0
    # three sweeps per pass are not offered
    assert lib.mxh_node_march(0, C.byref(lv), C.byref(bl), 1, v, Vecs(None, 512, None), v, 0, none, none, none,
                              0, 3) != 0


# A mapped right-hand side (b.map: vector v reads slot map[v] of b), as the
# solve's finest level reads the defects where they were formed: the same
# output, bit for bit, as the unmapped pass on the gathered right-hand sides.
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("mode", [10, 8, 5])
def test_afw_march_mapped_rhs(lib, prec, mode):
    n, h, rows, nv, nslot = 64, 1 / 64, 40, 3, 7
    rng = np.random.default_rng(5 + mode + prec)
    E, Ec = 2 * n * n, n * n // 2
    dt = cplx(prec)
    slots = (rng.standard_normal((nslot, E)) + 1j * rng.standard_normal((nslot, E))).astype(dt)
    vmap = np.array([5, 0, 3], np.int32)
    x = (rng.standard_normal((nv, E)) + 1j * rng.standard_normal((nv, E))).astype(dt)
    xc = (rng.standard_normal((nv, Ec)) + 1j * rng.standard_normal((nv, Ec))).astype(dt)
    dslots, dgath, dmap = Dev(lib, slots), Dev(lib, slots[vmap].copy()), Dev(lib, vmap)
    dx, dxc = Dev(lib, x), Dev(lib, xc)
    deta = Dev(lib, O.rod_eta(n).astype(np.float64 if prec == 0 else np.float32))
    lv = Level(n, n, h, h)
    bl = Bloch(np.cos(0.9), np.sin(0.9), np.cos(-0.4), np.sin(-0.4), None, 1, None)
    outs = []
    for mapped in (False, True):
        dxo = Dev(lib, np.zeros((nv, E), dt))
        dbc = Dev(lib, np.zeros((nv, Ec), dt))
        dxcz = Dev(lib, np.zeros((nv, Ec), dt))
        b = Vecs(dslots.p, E, dmap.p) if mapped else vecs(dgath, E)
        xc_in = vecs(dxc, Ec) if mode & 4 else NONE
        bc = vecs(dbc, Ec) if mode & 8 else NONE
        xcv = vecs(dxcz, Ec) if mode & 8 else NONE
        assert lib.mxh_afw_march(prec, C.byref(lv), C.byref(bl), nv, b, vecs(dx, E), vecs(dxo, E), deta.p, mode,
                                 xc_in, bc, xcv, rows, 2) == 0
        outs.append((dxo.get(), dbc.get()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
