"""ctypes views of the device C-ABI (include/mxh.h) for kernel-level GPU tests."""
import ctypes as C

import numpy as np


class Level(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("hx", C.c_double), ("hy", C.c_double)]


class Bloch(C.Structure):
    _fields_ = [("xr", C.c_double), ("xi", C.c_double), ("yr", C.c_double), ("yi", C.c_double),
                ("per_problem", C.c_void_p), ("vpp", C.c_int), ("vmap", C.c_void_p)]


class Vecs(C.Structure):
    _fields_ = [("p", C.c_void_p), ("stride", C.c_longlong), ("map", C.c_void_p)]


NONE = Vecs(None, 0, None)


def bind(L):
    """Declare the argument types of the launchers the tests call."""
    assert L.mxh_init() == 0
    L.mxh_malloc.restype = C.c_void_p
    L.mxh_malloc.argtypes = [C.c_size_t]
    L.mxh_free.argtypes = [C.c_void_p]
    L.mxh_h2d.argtypes = L.mxh_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.mxh_afw_march.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Bloch), C.c_int, Vecs, Vecs, Vecs, C.c_void_p,
                                C.c_int, Vecs, Vecs, Vecs, C.c_int, C.c_int]
    L.mxh_node_march.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Bloch), C.c_int, Vecs, Vecs, Vecs, C.c_int,
                                 Vecs, Vecs, Vecs, C.c_int, C.c_int]
    L.mxh_gram_work_bytes.restype = C.c_size_t
    L.mxh_gram_work_bytes.argtypes = [C.c_int, C.c_int]
    L.mxh_gram.argtypes = [C.POINTER(Level), C.POINTER(Bloch), C.c_int, C.c_void_p, C.c_int, Vecs, C.c_void_p,
                           C.c_void_p, C.c_void_p]
    L.mxh_gram_static.argtypes = [C.POINTER(Level), C.POINTER(Bloch), C.c_int, C.c_void_p, C.c_int, Vecs,
                                  C.c_void_p, C.c_void_p]
    L.mxh_gram_cols.argtypes = [C.POINTER(Level), C.POINTER(Bloch), C.c_int, C.c_void_p, C.c_void_p, C.c_int, Vecs,
                                C.c_void_p, C.c_void_p, C.c_void_p]
    L.mxh_error_string.restype = C.c_char_p
    L.mxh_error_string.argtypes = [C.c_int]
    L.mxh_defect.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Bloch), C.c_int, Vecs, Vecs, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_void_p]
    L.mxh_lift_rhs.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Bloch), C.c_int, Vecs, Vecs, Vecs, Vecs, Vecs,
                               C.c_void_p, C.c_void_p]
    L.mxh_lift_apply.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Bloch), C.c_int, Vecs, Vecs, Vecs, Vecs,
                                 C.c_double, C.c_void_p, C.c_void_p]
    L.mxh_coarse_setup.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.mxh_lobpcg_combine.argtypes = [C.c_longlong, C.c_int, C.c_void_p, C.c_int, Vecs, Vecs, C.c_void_p]
    L.mxh_combine_x.argtypes = [C.c_longlong, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, Vecs,
                                C.c_int, Vecs, C.c_int, C.c_void_p]
    return L


class Dev:
    """Device copy of a host array (freed with the object)."""

    def __init__(self, lib, a=None, nbytes=None):
        self.lib = lib
        self.a = None if a is None else np.ascontiguousarray(a)
        n = self.a.nbytes if a is not None else nbytes
        self.p = lib.mxh_malloc(n)
        assert self.p
        if a is not None:
            assert lib.mxh_h2d(self.p, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes) == 0

    def get(self, like=None):
        out = np.empty_like(self.a if like is None else like)
        assert self.lib.mxh_d2h(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes) == 0
        return out

    def __del__(self):
        self.lib.mxh_free(self.p)


def vecs(d, n):
    return Vecs(d.p if d is not None else None, n, None)
