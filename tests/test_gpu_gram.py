"""Kernel-level parity of the batched Gram products (mxh_gram) with the
oracle's energyprod/massprod (edge2d.c:1652-2027): for every problem of a
batch and every pair a >= b of its block, u_a^H K u_b (without the 1/(hx hy)
factor) and u_a^H M u_b, for block sizes that do and do not fill the kernel's
3 x 3 register blocks, with a problem map, on grids whose row bands hold one
or several rows."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_abi import Bloch, Dev, Level, bind, vecs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import maxwell_amd as M
    return bind(M.lib)


# (200, 12, 40): bands of 4 rows (2048 partial sets over 40 problems), so a
# thread carries x rows from one row of a segment column to the next, and
# 200 = 3 x 64 + 8 cells leaves a short last segment in every row.  The
# LOBPCG widths: 36 (3 x 12, the workloads' block) and 48 (3 x 16, the
# widest, 136 block pairs); 13 and 20 leave partial 3 x 3 register blocks;
# (60, 20, 40): bands of 2 rows over 40 problems
@pytest.mark.parametrize("n,m,nprob", [(32, 12, 3), (16, 1, 2), (64, 5, 1), (32, 24, 2), (8, 32, 1), (200, 12, 40),
                                       (24, 13, 3), (32, 36, 2), (16, 48, 1), (60, 20, 40)])
def test_gram_matches_oracle(lib, n, m, nprob):
    rng = np.random.default_rng(n * 100 + m)
    eta = O.rod_eta(n)
    E = 2 * n * n
    ks = [(0.3 * p - 0.2, 1.1 - 0.7 * p) for p in range(nprob)]
    U = rng.standard_normal((nprob * m, E)) + 1j * rng.standard_normal((nprob * m, E))
    # per-problem Bloch factors; the batch map reverses the problem order
    bf = np.array([[np.cos(kx), np.sin(kx), np.cos(ky), np.sin(ky)] for kx, ky in ks]).ravel()
    pmap = np.arange(nprob)[::-1].astype(np.int32).copy()
    dU, dbf, dmap, deta = Dev(lib, U), Dev(lib, bf), Dev(lib, pmap), Dev(lib, eta.astype(np.float64))
    np_ = m * (m + 1) // 2
    dout = Dev(lib, np.zeros(4 * np_ * nprob))
    dwork = Dev(lib, None, nbytes=lib.mxh_gram_work_bytes(m, nprob))
    lv = Level(n, n, 1.0 / n, 1.0 / n)
    bl = Bloch(1, 0, 1, 0, C.c_void_p(dbf.p), m, None)
    assert lib.mxh_gram(C.byref(lv), C.byref(bl), nprob, C.c_void_p(dmap.p), m, vecs(dU, E), C.c_void_p(deta.p),
                        C.c_void_p(dout.p), C.c_void_p(dwork.p)) == 0
    out = dout.get().reshape(nprob, np_, 4)
    for j in range(nprob):
        q = pmap[j]
        g = O.Grid2D(eta, kx=ks[q][0], ky=ks[q][1])
        Ub = U[q * m:(q + 1) * m]
        scale_k = 1.0 / (g.hx * g.hy)
        for a in range(m):
            for b in range(a + 1):
                p = a * (a + 1) // 2 + b
                refk = O.energyprod(g, 1.0, 0.0, Ub[a], Ub[b]) / scale_k
                refm = O.massprod(g, Ub[a], Ub[b])
                gk = out[j, p, 0] + 1j * out[j, p, 1]
                gm = out[j, p, 2] + 1j * out[j, p, 3]
                assert abs(gk - refk) <= 1e-12 * max(abs(refk), np.sqrt(abs(
                    O.energyprod(g, 1.0, 0.0, Ub[a], Ub[a]) * O.energyprod(g, 1.0, 0.0, Ub[b], Ub[b]))) / scale_k), \
                    (j, a, b, gk, refk)
                assert abs(gm - refm) <= 1e-12 * max(abs(refm), np.sqrt(abs(
                    O.massprod(g, Ub[a], Ub[a]) * O.massprod(g, Ub[b], Ub[b])))), (j, a, b, gm, refm)


# Column lists (mxh_gram_cols, the LOBPCG Ritz step's Gram over E, the W of
# unlocked vectors and P): per problem, the sums over the listed columns at
# list positions, against the full Gram's pairs of those columns.  Lists of
# one column, of every column, the workloads' [E W P] with some W and all P
# left out, and a ragged list that splits the register blocks; a reversed
# problem map (the lists are indexed by problem, the output by list position).
@pytest.mark.parametrize("n,m,nprob", [(32, 36, 4), (24, 13, 3), (200, 36, 2)])
def test_gram_cols_matches_full_gram(lib, n, m, nprob):
    rng = np.random.default_rng(7 * n + m)
    eta = O.rod_eta(n)
    E = 2 * n * n
    U = rng.standard_normal((nprob * m, E)) + 1j * rng.standard_normal((nprob * m, E))
    ks = [(0.25 * p - 0.3, 0.9 - 0.5 * p) for p in range(nprob)]
    bf = np.array([[np.cos(kx), np.sin(kx), np.cos(ky), np.sin(ky)] for kx, ky in ks]).ravel()
    pmap = np.arange(nprob)[::-1].astype(np.int32).copy()
    b = m // 3
    lists = [[0], list(range(m)), list(range(b)) + [b + i for i in range(b) if i % 3 != 1],
             sorted(rng.choice(m, size=m // 2 + 1, replace=False).tolist())]
    cmap = np.zeros((nprob, m + 1), np.int32)
    for q in range(nprob):
        cl = lists[q % len(lists)]
        cmap[q, 0] = len(cl)
        cmap[q, 1:1 + len(cl)] = cl
    dU, dbf, dmap, deta = Dev(lib, U), Dev(lib, bf), Dev(lib, pmap), Dev(lib, eta.astype(np.float64))
    dcols = Dev(lib, cmap)
    np_ = m * (m + 1) // 2
    dfull = Dev(lib, np.zeros(4 * np_ * nprob))
    dsub = Dev(lib, np.full(4 * np_ * nprob, np.nan))
    dwork = Dev(lib, None, nbytes=lib.mxh_gram_work_bytes(m, nprob))
    lv = Level(n, n, 1.0 / n, 1.0 / n)
    bl = Bloch(1, 0, 1, 0, C.c_void_p(dbf.p), m, None)
    args = (C.byref(lv), C.byref(bl), nprob, C.c_void_p(dmap.p))
    assert lib.mxh_gram(*args, m, vecs(dU, E), C.c_void_p(deta.p), C.c_void_p(dfull.p), C.c_void_p(dwork.p)) == 0
    assert lib.mxh_gram_cols(*args, C.c_void_p(dcols.p), m, vecs(dU, E), C.c_void_p(deta.p), C.c_void_p(dsub.p),
                             C.c_void_p(dwork.p)) == 0
    full = dfull.get().reshape(nprob, np_, 4)
    sub = dsub.get().reshape(nprob, np_, 4)
    for j in range(nprob):
        q = pmap[j]
        cl = cmap[q, 1:1 + cmap[q, 0]]
        diag = np.array([full[j, c * (c + 1) // 2 + c] for c in range(m)])
        for a in range(len(cl)):
            for bb in range(a + 1):
                ca, cb = cl[a], cl[bb]
                ref = full[j, ca * (ca + 1) // 2 + cb]
                got = sub[j, a * (a + 1) // 2 + bb]
                # K and M parts each to rounding of the Cauchy-Schwarz bound
                for part in (0, 2):
                    scale = np.sqrt(abs(diag[ca, part] * diag[cb, part]))
                    assert abs(got[part] - ref[part]) <= 1e-13 * scale, (j, a, bb, part)
                    assert abs(got[part + 1] - ref[part + 1]) <= 1e-13 * scale, (j, a, bb, part)
