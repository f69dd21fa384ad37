"""Householder M-orthonormalisation of the module API (orthonormalize_edge2d,
edge2d.c:6154-6296) on the device -- one launch sequence over a contiguous
copy of the k vectors, no host round trip per entry -- against the oracle's
restatement of the same reflections (tests/test_oracle.py pins it by
M-orthonormality, nested spans, the Cholesky-QR basis and the unit-vector
closed form): outputs and tau to 1e-12, for a few block sizes (the drivers'
Ritz extrapolation orthonormalises up to 3 x (eigs + throwaway) vectors),
rectangular grids and complex Bloch factors."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,k,kx,ky", [(16, 16, 12, 0.3, -1.1), (32, 24, 36, np.pi, 0.5), (8, 12, 5, 0.0, 0.0),
                                           (64, 64, 24, -2.0, 2.9)])
def test_orthonormalize_matches_oracle(nx, ny, k, kx, ky):
    import maxwell_amd as M
    print("orthonormalize %dx%d k %d" % (nx, ny, k), flush=True)
    rng = np.random.default_rng(nx + k)
    eta = 1.0 / rng.uniform(1.0, 12.0, size=(ny, nx))
    hx, hy = 1.0 / nx, 0.8 / ny
    G = M.Grid(eta, kx=kx, ky=ky, hx=hx, hy=hy)
    xf, yf = G.factors
    g = O.Grid2D(eta, hx, hy)
    g.set_factors(xf, yf)
    X = rng.standard_normal((k, 2 * nx * ny)) + 1j * rng.standard_normal((k, 2 * nx * ny))
    edges = [M.Edge(G, X[i]) for i in range(k)]
    work = M.Edge(G)
    ptrs = (C.c_void_p * k)(*[e.h.value for e in edges])
    tau = (M.Cplx * k)()
    M.lib.orthonormalize_edge2d(k, ptrs, tau, work.h)
    out = np.array([e.get() for e in edges])
    ref, tref, _ = O.orthonormalize(g, X)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    terr = max(abs(tau[i].value() - tref[i]) / abs(tref[i]) for i in range(k))
    print("    outputs rel err %.2e, tau rel err %.2e" % (err, terr))
    assert err < 1e-12 and terr < 1e-12
    assert np.abs(work.get() - ref[0]).max() < 1e-12 * np.abs(ref).max()
    Mm = O.dense_edge(g, 0, 1) if nx * ny <= 1024 else None
    if Mm is not None:
        assert np.abs(out.conj() @ Mm @ out.T - np.eye(k)).max() < 1e-12
