"""Kernel-level parity of the square-tile smoothers (mxh_afw_tile,
mxh_node_tile), one or two sweeps per launch, with the oracle's four-colour
sweeps (edge2d.c:2656-3042, node2d.c:629-1138), transfers (edge2d.c:5583-5705,
node2d.c:1140-1273) and residuals, called through the device C-ABI
(include/mxh.h).  Grids span several tiles in both directions or sit inside
one tile several periods wide, and wrap with complex Bloch factors; two sweeps
per launch run with the residual and its restriction (forward) or after the
prolongation (backward)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_abi import NONE, Bloch, Dev, Level, Vecs, bind, vecs
from test_gpu_march import close, cplx, tolerance

pytestmark = pytest.mark.gpu

# (grid n, mesh width h, kx, ky): 256 spans 6 x 6 FP32 tiles and 10 x 10 FP64
# tiles; the small grids sit inside one tile
CASES = [(256, 1 / 16, 0.7, -1.3), (64, 1 / 64, np.pi, 0.4), (16, 1 / 16, -2.0, 1.1), (8, 1 / 8, 0.0, 0.0)]
MODES = [0, 1, 2, 5, 8, 10]


@pytest.fixture(scope="module")
def lib():
    import maxwell_amd as M
    L = bind(M.lib)
    L.mxh_afw_tile.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Bloch), C.c_int, Vecs, Vecs, Vecs, C.c_void_p,
                               C.c_int, C.c_int, Vecs, Vecs, Vecs]
    L.mxh_node_tile.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Bloch), C.c_int, Vecs, Vecs, Vecs, C.c_int,
                                C.c_int, Vecs, Vecs, Vecs]
    return L


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,h,kx,ky", CASES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("nsweep", [1, 2])
def test_afw_tile_matches_oracle(lib, prec, n, h, kx, ky, mode, nsweep):
    print("afw_tile prec %d n %d mode %d nsweep %d" % (prec, n, mode, nsweep), flush=True)
    rng = np.random.default_rng(3 * n + mode + 100 * nsweep)
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
    xc_in = vecs(dxc, Ec) if mode & 4 else NONE
    bc = vecs(dbc, Ec) if mode & 8 else NONE
    xcv = vecs(dxc, Ec) if mode & 8 else NONE
    assert lib.mxh_afw_tile(prec, C.byref(lv), C.byref(bl), nv, vecs(db, E), vecs(dx, E), vecs(dxo, E), deta.p,
                            nsweep, mode, xc_in, bc, xcv) == 0
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
            xs = ref if prec == 0 else got[v].astype(complex)
            bb = b[v].astype(dt).astype(complex)
            rc = O.restrict_edge(g, O.addeval(g, -1.0, -1.0, xs, bb))
            scale = np.abs(rc).max() if prec == 0 else np.abs(O.restrict_edge(g, O.addeval(g, 1.0, 1.0, xs))).max()
            err = np.abs(dbc.get()[v] - rc).max() / scale
            print("    restricted residual rel err %.2e" % err)
            assert err <= tolerance(prec, h)
            assert np.abs(dxc.get()[v]).max() == 0


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,h,kx,ky", CASES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("nsweep", [1, 2])
def test_node_tile_matches_oracle(lib, prec, n, h, kx, ky, mode, nsweep):
    print("node_tile prec %d n %d mode %d nsweep %d" % (prec, n, mode, nsweep), flush=True)
    rng = np.random.default_rng(5 * n + mode + 100 * nsweep)
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
    xc_in = vecs(dxc, Nc) if mode & 4 else NONE
    bc = vecs(dbc, Nc) if mode & 8 else NONE
    xcv = vecs(dxc, Nc) if mode & 8 else NONE
    assert lib.mxh_node_tile(prec, C.byref(lv), C.byref(bl), nv, vecs(db, N), vecs(dx, N), vecs(dxo, N), nsweep,
                             mode, xc_in, bc, xcv) == 0
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


def test_tile_rejects_sweeps_beyond_the_halo(lib):
    lv = Level(16, 16, 1 / 16, 1 / 16)
    bl = Bloch(1, 0, 1, 0, None, 1, None)
    d = Dev(lib, np.zeros(4 * 2 * 256, np.complex128))
    b, x, xo = Vecs(d.p, 512, None), Vecs(d.p + 16 * 512, 512, None), Vecs(d.p + 32 * 512, 512, None)
    c = Vecs(d.p + 48 * 512, 128, None)
    # three sweeps, and two sweeps backward with a residual, exceed the halo
    assert lib.mxh_afw_tile(0, C.byref(lv), C.byref(bl), 1, b, x, xo, d.p, 3, 0, NONE, NONE, NONE) != 0
    assert lib.mxh_afw_tile(0, C.byref(lv), C.byref(bl), 1, b, x, xo, d.p, 2, 9, NONE, c, c) != 0
    assert lib.mxh_node_tile(0, C.byref(lv), C.byref(bl), 1, b, x, xo, 3, 0, NONE, NONE, NONE) != 0
