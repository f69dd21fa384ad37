"""Kernel-level parity of the streaming kernels of a block step (mxh_defect,
mxh_lift_rhs, mxh_lift_apply) with the oracle's operators, called through the
device C-ABI (include/mxh.h).

The grids are rectangular with hx != hy and complex Bloch factors.  Rows of
up to 520 give every row-band several rows (a band is ceil(ny / 128) rows),
so the column walk through a band, the empty last bands, the interior/frame
split and the wrap are all exercised.  Reference operators: the PINVIT defect
(K + (1 - lam) M) e of exp_bloch2d.c:897-920 is the oracle's addeval; the
lifting right-hand side G^H M (e - x) and update u - x + alpha G xg of
edge2d.c:6319-6327 are the oracle's adjgradient, addeval and gradient.  The
plane-wave deflation is pinned by what the mathematics fixes: phi is an
eigenvector of the node operator G^H M G (checked here), so the deflated right
hand side must be orthogonal to phi and coef_v must be phi's coefficient in
the solution of the node problem."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_abi import NONE, Bloch, Dev, Level, bind, vecs

pytestmark = pytest.mark.gpu

# (nx, ny, hx, hy, kx, ky)
CASES = [(40, 300, 1 / 40, 1 / 200, 0.7, -1.3), (136, 520, 1 / 136, 1 / 520, np.pi, 0.4),
         (24, 24, 1 / 24, 1 / 16, -2.0, 1.1), (8, 8, 1 / 8, 1 / 8, 0.3, 0.9)]
NV = 3


@pytest.fixture(scope="module")
def lib():
    import maxwell_amd as M
    return bind(M.lib)


def cplx(prec):
    return np.complex128 if prec == 0 else np.complex64


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(nx, ny, hx, hy, kx, ky, seed):
    rng = np.random.default_rng(seed)
    eta = 1.0 / rng.uniform(1.0, 12.0, size=(ny, nx))
    g = O.Grid2D(eta, hx=hx, hy=hy, kx=kx, ky=ky)
    lv = Level(nx, ny, hx, hy)
    bl = Bloch(np.cos(kx), np.sin(kx), np.cos(ky), np.sin(ky), None, 1, None)
    return rng, eta, g, lv, bl


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def plane_wave(g):
    """Lowest Bloch plane wave of the node space, phi(i + nx, j) = xf phi(i, j)."""
    tx, ty = np.angle(g.xf) / g.nx, np.angle(g.yf) / g.ny
    i, j = np.meshgrid(np.arange(g.nx), np.arange(g.ny))
    return np.exp(1j * (tx * i + ty * j)).ravel()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nx,ny,hx,hy,kx,ky", CASES)
def test_defect_matches_oracle(lib, prec, nx, ny, hx, hy, kx, ky):
    print("defect prec %d grid %dx%d" % (prec, nx, ny), flush=True)
    rng, eta, g, lv, bl = setup(nx, ny, hx, hy, kx, ky, nx + ny)
    E = 2 * nx * ny
    e = crandn(rng, NV, E)
    lam = rng.uniform(0.5, 40.0, NV)
    de, dr = Dev(lib, e), Dev(lib, np.zeros((NV, E), cplx(prec)))
    deta, dlam = Dev(lib, eta), Dev(lib, lam)
    dout = Dev(lib, np.zeros(2 * NV))
    dwork = Dev(lib, nbytes=8 * 2 * NV * 256)
    assert lib.mxh_defect(prec, C.byref(lv), C.byref(bl), NV, vecs(de, E), vecs(dr, E), deta.p, dlam.p, dout.p,
                          dwork.p) == 0
    assert lib.mxh_sync() == 0
    r, out = dr.get(), dout.get()
    for v in range(NV):
        ref = O.addeval(g, 1.0, 1.0 - lam[v], e[v])
        me = O.addeval(g, 0.0, 1.0, e[v])
        err = rel(r[v], ref)
        print("    v %d r rel err %.2e" % (v, err))
        assert err <= (1e-13 if prec == 0 else 2e-7)
        # the norms are summed in FP64 before r is rounded to its precision
        assert abs(out[2 * v] - np.vdot(ref, ref).real) <= 1e-12 * np.vdot(ref, ref).real
        assert abs(out[2 * v + 1] - np.vdot(me, me).real) <= 1e-12 * np.vdot(me, me).real


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("deflate", [False, True])
@pytest.mark.parametrize("with_u", [False, True])
@pytest.mark.parametrize("nx,ny,hx,hy,kx,ky", CASES)
def test_lift_rhs_matches_oracle(lib, prec, deflate, with_u, nx, ny, hx, hy, kx, ky):
    """with_u: the kernel also stores u = e - x (FP64) in a third buffer."""
    print("lift_rhs prec %d deflate %d u %d grid %dx%d" % (prec, deflate, with_u, nx, ny), flush=True)
    rng, eta, g, lv, bl = setup(nx, ny, hx, hy, kx, ky, 3 * nx + ny)
    E, N = 2 * nx * ny, nx * ny
    e = crandn(rng, NV, E)
    x = crandn(rng, NV, E).astype(cplx(prec))
    de, dx = Dev(lib, e), Dev(lib, x)
    du = Dev(lib, np.full((NV, E), np.nan + 0j)) if with_u else None
    dbg = Dev(lib, np.zeros((NV, N), cplx(prec)))
    dxg = Dev(lib, crandn(rng, NV, N).astype(cplx(prec)))
    dcoef = Dev(lib, np.zeros(2 * NV))
    dwork = Dev(lib, nbytes=8 * 2 * 513 * NV)
    assert lib.mxh_lift_rhs(prec, C.byref(lv), C.byref(bl), NV, vecs(de, E), vecs(dx, E),
                            vecs(du, E) if with_u else NONE, vecs(dbg, N), vecs(dxg, N),
                            dcoef.p if deflate else None, dwork.p if deflate else None) == 0
    assert lib.mxh_sync() == 0
    bg, xg, coef = dbg.get(), dxg.get(), dcoef.get()
    assert not xg.any()
    if with_u:
        # e - x in FP64, every entry (the buffer started as NaN)
        assert np.array_equal(du.get(), e - x.astype(np.complex128))
    phi = plane_wave(g)
    A = O.addeval_node(g, 1.0, phi)
    mu = np.vdot(phi, A).real / N
    # mu is small against the stencil's coefficients on these meshes (1e-4 of
    # them on the tall grids), so both the oracle's mu and the device's lose
    # digits in proportion
    assert rel(A, mu * phi) <= 1e-9
    for v in range(NV):
        d = O.adjgradient(g, O.addeval(g, 0.0, 1.0, e[v] - x[v].astype(np.complex128)))
        ref = d
        if deflate:
            a = np.vdot(phi, d) / N
            ref = d - a * phi
            assert abs(np.vdot(phi, bg[v].astype(np.complex128))) <= (1e-11 if prec == 0 else 1e-5) * np.abs(d).max() * N
            c = complex(coef[2 * v], coef[2 * v + 1])
            print("    v %d coef %s expected %s" % (v, c, a / mu))
            # FP32: the wrapped reads of x round in FP32 (about 1e-7 of |d| at
            # the frame), a small part of the small phi^H d
            assert abs(c - a / mu) <= (1e-8 if prec == 0 else 1e-5) * abs(a / mu)
        err = rel(bg[v], ref)
        print("    v %d bg rel err %.2e" % (v, err))
        # FP32: x and the Bloch factors of the wrapped reads are rounded
        assert err <= (1e-13 if prec == 0 else 1e-6)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("nx,ny,hx,hy,kx,ky", CASES)
def test_lift_apply_matches_oracle(lib, prec, inplace, nx, ny, hx, hy, kx, ky):
    print("lift_apply prec %d inplace %d grid %dx%d" % (prec, inplace, nx, ny), flush=True)
    rng, eta, g, lv, bl = setup(nx, ny, hx, hy, kx, ky, 5 * nx + ny)
    E, N = 2 * nx * ny, nx * ny
    u = crandn(rng, NV, E)
    x = crandn(rng, NV, E).astype(cplx(prec))
    xg = crandn(rng, NV, N).astype(cplx(prec))
    coef = rng.standard_normal(2 * NV)
    alpha = -0.75
    du, dx, dxg, dcoef = Dev(lib, u), Dev(lib, x), Dev(lib, xg), Dev(lib, coef)
    de = du if inplace else Dev(lib, np.zeros((NV, E), np.complex128))
    dwork = Dev(lib, nbytes=8 * 2 * 513 * NV)
    assert lib.mxh_lift_apply(prec, C.byref(lv), C.byref(bl), NV, vecs(dxg, N), vecs(du, E), vecs(dx, E),
                              vecs(de, E), alpha, dcoef.p, dwork.p) == 0
    assert lib.mxh_sync() == 0
    out = de.get(u)
    phi = plane_wave(g)
    for v in range(NV):
        xgv = xg[v].astype(np.complex128)
        a = complex(coef[2 * v], coef[2 * v + 1]) - np.vdot(phi, xgv) / N
        ref = O.gradient(g, xgv + a * phi, u[v] - x[v].astype(np.complex128), alpha=alpha)
        err = rel(out[v], ref)
        print("    v %d e rel err %.2e" % (v, err))
        assert err <= (1e-13 if prec == 0 else 1e-6)
