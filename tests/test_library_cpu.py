"""CPU checks of the C-ABI library: it loads without a GPU and exports every
function its headers declare (no compute calls)."""
import ctypes
import os

import maxwell_amd as M


def test_library_loads_and_exports_every_declared_symbol():
    names = M.header_functions()
    assert len(names) > 80
    missing = [n for n in names if not hasattr(M.lib, n)]
    assert not missing, missing


def test_reference_api_entry_points_present():
    # the calls the reference drivers make (exp_bloch2d.c:335-1261, exp_blochsurf.c:520-1487)
    for n in ["new_grid2d", "coarsen_grid2d", "setpattern_grid2d", "new_edge2d", "new_node2d", "new_matrix",
              "densematrix_edge2d", "densematrix_node2d", "potrf_matrix", "sygv_matrix", "syev_matrix",
              "denseto_edge2d", "zero_edge2d", "copy_edge2d", "swap_edge2d", "scale_edge2d", "add_edge2d",
              "dotprod_edge2d", "norm2_edge2d", "addeval_edge2d", "energyprod_edge2d", "massprod_edge2d",
              "prolongation_edge2d", "vcycle_edge2d", "vcycle_node2d", "adjgradient_node2d", "gradient_node2d",
              "orthonormalize_edge2d", "pinvit_edge2d", "reduce_max_real"]:
        assert hasattr(M.lib, n), n


def test_grid_struct_layout_matches_header():
    # offsets of the public grid2d fields (include/grid2d.h)
    G = M.Grid2DStruct
    assert G.hx.offset == 0 and G.nx.offset == 16 and G.xfactor.offset == 40
    assert G.yfactor.offset == 56 and G.eps.offset == 72 and ctypes.sizeof(G) == 88


def test_dense_linalg_without_gpu():
    """The host dense routines (linalg.h) need no GPU."""
    import numpy as np
    import scipy.linalg as sl
    n = 10
    rng = np.random.default_rng(0)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = X @ X.conj().T + n * np.eye(n)
    Y = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = Y + Y.conj().T
    ma, mb = M.lib.new_matrix(n, n), M.lib.new_matrix(n, n)
    Af, Bf = np.asfortranarray(A), np.asfortranarray(B)  # keep alive across the copies
    ctypes.memmove(ma.contents.a, Af.ctypes.data, Af.nbytes)
    ctypes.memmove(mb.contents.a, Bf.ctypes.data, Bf.nbytes)
    w = np.zeros(n)
    assert M.lib.sygv_matrix(ma, mb, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    assert np.allclose(w, sl.eigh(A, B, eigvals_only=True), rtol=1e-12, atol=1e-12)
    V = np.ctypeslib.as_array(ctypes.cast(ma.contents.a, ctypes.POINTER(ctypes.c_double)), shape=(2 * n * n,))
    V = V.view(np.complex128).reshape(n, n).T
    assert np.allclose(V.conj().T @ B @ V, np.eye(n), atol=1e-12)
    M.lib.del_matrix(ma)
    M.lib.del_matrix(mb)
