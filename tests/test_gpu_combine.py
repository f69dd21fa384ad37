"""Kernel-level checks of the block updates against numpy: the LOBPCG update
(mxh_lobpcg_combine: P_new = [W P] C_WP, E_new = E C_E + P_new into the E and
P slots of each problem's [E | W | P] block, W untouched) with a problem map,
including coefficient rows that the Ritz step zeroes for the columns it drops
(those inputs are not read: NaN there must not reach the outputs), and the
general combine (mxh_combine_x: out_j = sum_a C[a, j] in_a, outputs from
nsplit on shifted by gap)."""
import ctypes as C

import numpy as np
import pytest

from gpu_abi import Dev, bind, vecs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import maxwell_amd as M
    return bind(M.lib)


def cplx(r, shape):
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


@pytest.mark.parametrize("m", [2, 8, 12, 16])
def test_lobpcg_combine_matches_numpy(lib, m):
    r = np.random.default_rng(m)
    n, nprob = 1000, 3
    k = 3 * m
    X = cplx(r, (nprob, k, n))
    C_ = cplx(r, (nprob, k, 2 * m))          # column-major 3m x 2m per problem
    # the Ritz step's dropped columns: a single W column (the partner of a
    # kept one) of problem 0, a W pair and the whole P block of problem 1,
    # the last E pair of problem 2; their inputs hold NaN
    drops = {0: [m + 1], 1: list(range(m, m + 2)) + list(range(2 * m, 3 * m)), 2: [m - 2, m - 1]}
    for q, cols in drops.items():
        C_[q, cols, :] = 0.0
        X[q, cols, :] = np.nan
    pmap = np.array([2, 0, 1], np.int32)    # batch slot j -> problem pmap[j]
    Cb = np.stack([C_[pmap[j]] for j in range(nprob)])
    dX = Dev(lib, X.copy())
    dC = Dev(lib, np.ascontiguousarray(Cb.transpose(0, 2, 1)))  # a + q k layout
    dmap = Dev(lib, pmap)
    assert lib.mxh_lobpcg_combine(n, nprob, C.c_void_p(dmap.p), m, vecs(dX, n), vecs(dX, n),
                                  C.c_void_p(dC.p)) == 0
    out = dX.get()
    for q in range(nprob):
        Xq = np.nan_to_num(X[q], nan=0.0)
        Ce = C_[q][:, :m]
        P_new = Ce[m:].T @ Xq[m:]
        E_new = Ce[:m].T @ Xq[:m] + P_new
        assert np.abs(out[q, :m] - E_new).max() < 1e-12 * np.abs(E_new).max()
        assert np.abs(out[q, 2 * m:] - P_new).max() < 1e-12 * max(np.abs(P_new).max(), 1e-300) or \
            np.abs(P_new).max() == 0.0 and np.abs(out[q, 2 * m:]).max() == 0.0
        W = X[q, m:2 * m]
        assert np.array_equal(np.isnan(out[q, m:2 * m]), np.isnan(W))
        assert np.array_equal(out[q, m:2 * m][~np.isnan(W)], W[~np.isnan(W)])


@pytest.mark.parametrize("nin,nout,nsplit,gap", [(8, 6, 6, 0), (36, 12, 12, 0), (48, 32, 20, 4), (5, 3, 1, 2)])
def test_combine_x_matches_numpy(lib, nin, nout, nsplit, gap):
    r = np.random.default_rng(nin + nout)
    n, nprob = 777, 2
    yvpp = nout + gap
    X = cplx(r, (nprob, nin, n))
    C_ = cplx(r, (nprob, nin, nout))
    Y0 = cplx(r, (nprob, yvpp, n))
    dX, dY = Dev(lib, X), Dev(lib, Y0.copy())
    dC = Dev(lib, np.ascontiguousarray(C_.transpose(0, 2, 1)))
    assert lib.mxh_combine_x(n, nprob, None, nin, nout, nsplit, gap, vecs(dX, n), nin, vecs(dY, n), yvpp,
                             C.c_void_p(dC.p)) == 0
    Y = dY.get()
    for q in range(nprob):
        ref = C_[q].T @ X[q]
        slots = [j if j < nsplit else j + gap for j in range(nout)]
        assert np.abs(Y[q, slots] - ref).max() < 1e-12 * np.abs(ref).max()
        untouched = [s for s in range(yvpp) if s not in slots]
        assert np.array_equal(Y[q, untouched], Y0[q, untouched])
