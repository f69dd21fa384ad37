"""The Gram entry point that allocates nothing (mxh_gram_static, partial sums
in the library's static device array) against the workspace entry point
mxh_gram it is a variant of, through the device C-ABI: two calls in a row,
the second on a smaller input (the array keeps the first call's partial sums,
which must not leak into the second), a block size whose partial sums are
split into fewer row bands to fit the array, and a request beyond the array,
which must be refused without a launch."""
import ctypes as C

import numpy as np
import pytest

from gpu_abi import Bloch, Dev, Level, bind, vecs

pytestmark = pytest.mark.gpu

ESCRATCH = -2
SCRATCH_BYTES = 8 << 20


@pytest.fixture(scope="module")
def lib():
    import maxwell_amd as M
    return bind(M.lib)


def gram_pair(lib, nx, ny, nprob, m, seed):
    """(static result, workspace result) for nprob random problems of m vectors
    on an nx x ny grid with a per-problem Bloch table and a problem map."""
    print("gram nx %d ny %d nprob %d m %d" % (nx, ny, nprob, m), flush=True)
    rng = np.random.default_rng(seed)
    n = 2 * nx * ny
    np_ = m * (m + 1) // 2
    # one spare problem, so the map can skip one
    u = rng.standard_normal((nprob + 1, m, n)) + 1j * rng.standard_normal((nprob + 1, m, n))
    eta = 1.0 / rng.uniform(1.0, 12.0, size=(ny, nx))
    k = rng.uniform(-np.pi, np.pi, size=(nprob + 1, 2))
    table = np.stack([np.cos(k[:, 0]), np.sin(k[:, 0]), np.cos(k[:, 1]), np.sin(k[:, 1])], axis=1).ravel()
    pmap = np.array([p if p < nprob // 2 else p + 1 for p in range(nprob)], np.int32)
    du, deta, dtab, dmap = Dev(lib, u), Dev(lib, eta), Dev(lib, table), Dev(lib, pmap)
    lv = Level(nx, ny, 1.0 / nx, 1.0 / ny)
    bl = Bloch(1.0, 0.0, 1.0, 0.0, dtab.p, m, None)
    out_s = Dev(lib, np.full(4 * np_ * nprob, np.nan))
    out_w = Dev(lib, np.full(4 * np_ * nprob, np.nan))
    work = Dev(lib, nbytes=lib.mxh_gram_work_bytes(m, nprob))
    rc = lib.mxh_gram_static(C.byref(lv), C.byref(bl), nprob, dmap.p, m, vecs(du, n), deta.p, out_s.p)
    assert rc == 0, lib.mxh_error_string(rc)
    assert lib.mxh_gram(C.byref(lv), C.byref(bl), nprob, dmap.p, m, vecs(du, n), deta.p, out_w.p, work.p) == 0
    assert lib.mxh_sync() == 0
    return out_s.get(), out_w.get()


def test_static_gram_matches_workspace_gram_twice_then_smaller(lib):
    # block 12 (78 pairs): the static array holds all 2048 partial sets, the
    # launch is the workspace one's, so the sums agree bit for bit
    s, w = gram_pair(lib, 96, 80, 40, 12, 1)
    assert np.isfinite(w).all() and np.array_equal(s, w)
    # a second, smaller call right after it
    s, w = gram_pair(lib, 40, 24, 3, 5, 2)
    assert np.isfinite(w).all() and np.array_equal(s, w)


def test_static_gram_with_fewer_bands_agrees_to_rounding(lib):
    # block 24 (300 pairs): 873 partial sets fit the array against the
    # workspace's 2048, so the bands and the order of the final sums differ
    s, w = gram_pair(lib, 64, 64, 16, 24, 3)
    assert np.isfinite(s).all()
    assert np.abs(s - w).max() <= 1e-13 * np.abs(w).max()
    # and a smaller block again on the same array
    s, w = gram_pair(lib, 32, 48, 5, 7, 4)
    assert np.array_equal(s, w)


def test_static_gram_refuses_requests_beyond_the_array(lib):
    m, nx, ny = 48, 8, 8
    np_ = m * (m + 1) // 2
    cap = SCRATCH_BYTES // (32 * np_)
    nprob = cap + 1
    n = 2 * nx * ny
    # buffers of the full request, so that even a wrong launch stays in bounds
    du = Dev(lib, np.zeros((nprob, m, n), np.complex128))
    deta = Dev(lib, np.ones((ny, nx)))
    out = Dev(lib, np.full(4 * np_ * nprob, 7.0))
    lv = Level(nx, ny, 1.0 / nx, 1.0 / ny)
    bl = Bloch(1.0, 0.0, 1.0, 0.0, None, m, None)
    rc = lib.mxh_gram_static(C.byref(lv), C.byref(bl), nprob, None, m, vecs(du, n), deta.p, out.p)
    assert rc == ESCRATCH
    assert b"static scratch" in lib.mxh_error_string(rc)
    assert lib.mxh_sync() == 0
    assert (out.get() == 7.0).all()
    # one problem fewer fits
    rc = lib.mxh_gram_static(C.byref(lv), C.byref(bl), cap, None, m, vecs(du, n), deta.p, out.p)
    assert rc == 0 and lib.mxh_sync() == 0
    assert np.array_equal(out.get()[:4 * np_ * cap], np.zeros(4 * np_ * cap))
