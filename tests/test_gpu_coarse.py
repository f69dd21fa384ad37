"""Per-k coarse set-up on the device (mxh_coarse_setup): the Cholesky factors of
the coarse edge matrix K + M and of the coarse node Laplacian (with the
Gamma-point fill of node2d.c:1454-1458), assembled and factored by one
workgroup per problem, against numpy's Cholesky of the oracle's dense
matrices (edge2d.c:5870-6110, node2d.c:1429-1602) for a batch of Bloch
parameters that includes Gamma and the zone corner."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_abi import Dev, bind

pytestmark = pytest.mark.gpu

KS = [(0.0, 0.0), (0.3, -1.1), (np.pi, np.pi), (np.pi, 0.0), (2.0, 0.7), (-0.4, 3.0)]


@pytest.fixture(scope="module")
def L():
    import maxwell_amd
    return bind(maxwell_amd.lib)


@pytest.mark.parametrize("n0", [2, 4])
def test_coarse_factors_match_oracle(L, n0):
    rng = np.random.default_rng(n0)
    eta = 1.0 / rng.uniform(1.0, 12.0, size=(n0, n0))
    P, ne, nn = len(KS), 2 * n0 * n0, n0 * n0
    bloch = np.array([[np.cos(a), np.sin(a), np.cos(b), np.sin(b)] for a, b in KS]).ravel()
    db, de = Dev(L, bloch), Dev(L, eta.ravel())
    Ae = Dev(L, nbytes=16 * ne * ne * P)
    An = Dev(L, nbytes=16 * nn * nn * P)
    info = Dev(L, np.full(P, -1, np.int32))
    print("coarse setup n0 =", n0, flush=True)
    assert L.mxh_coarse_setup(n0, P, db.p, de.p, Ae.p, An.p, info.p) == 0
    assert (info.get() == 0).all()
    ae = Ae.get(np.zeros(ne * ne * P, complex)).reshape(P, ne, ne)
    an = An.get(np.zeros(nn * nn * P, complex)).reshape(P, nn, nn)
    for p, (kx, ky) in enumerate(KS):
        g = O.Grid2D(eta, kx=kx, ky=ky)
        g.set_factors(complex(bloch[4 * p], bloch[4 * p + 1]), complex(bloch[4 * p + 2], bloch[4 * p + 3]))
        ref_e = np.linalg.cholesky(O.dense_edge(g, 1.0, 1.0))
        ref_n = np.linalg.cholesky(O.dense_node(g, fixnull=True))
        got_e, got_n = ae[p].T, an[p].T  # column-major on the device
        assert np.abs(got_e - ref_e).max() <= 1e-13 * np.abs(ref_e).max(), (n0, p)
        assert np.abs(got_n - ref_n).max() <= 1e-12 * np.abs(ref_n).max(), (n0, p)


def test_coarse_setup_rejects_large_coarse_grids(L):
    """2 n0^2 above the LDS limit (64) is refused: the solver assembles those
    on the host."""
    assert L.mxh_coarse_setup(6, 1, None, None, None, None, None) != 0
