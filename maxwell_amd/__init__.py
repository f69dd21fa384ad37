"""maxwell_amd -- MI355X (gfx950) Bloch eigensolver behind the reference's C module API.

The product is the C-ABI shared library ``maxwell_amd/libmaxwell_amd.so`` (headers in
``include/``).  This module only loads it with ctypes for Python callers (tests,
benchmarks); it has no compute path of its own and raises if the library is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libmaxwell_amd.so")
INCLUDE = os.path.join(os.path.dirname(_HERE), "include")

MX_PREC_F64 = 0
MX_PREC_MIXED = 1
MX_METHOD_PINVIT = 0
MX_METHOD_LOBPCG = 1
MX_ERRORS = {0: "ok", 1: "invalid argument", 2: "out of memory", 3: "HIP error",
             4: "Ritz eigenproblem failed", 5: "non-finite residual"}


def build(verbose=False):
    """Compile the HIP kernels for gfx950 and the C host layer (in-tree)."""
    cmd = ["make", "-C", os.path.join(_HERE, "csrc"), "-B", "-j8"]
    out = None if verbose else subprocess.DEVNULL
    subprocess.check_call(cmd, stdout=out)


def load():
    if not os.path.exists(LIBPATH):
        raise RuntimeError("maxwell_amd: %s is not built; run __graft_entry__.build()" % LIBPATH)
    return C.CDLL(LIBPATH)


lib = load()
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class Cplx(C.Structure):
    """ABI twin of C99 `double complex` (two doubles, passed in SSE registers)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)

    def value(self):
        return complex(self.re, self.im)


class MxConfig(C.Structure):
    _fields_ = [("n0", C.c_int), ("levels", C.c_int), ("precision", C.c_int), ("block", C.c_int),
                ("max_problems", C.c_int), ("smoother_steps", C.c_int), ("prec_steps", C.c_int),
                ("gradient_steps", C.c_int), ("phase_kernels", C.c_int), ("lock", C.c_int),
                ("march_min", C.c_int), ("march_rows", C.c_int), ("march_fuse", C.c_int), ("autotune", C.c_int),
                ("graphs", C.c_int), ("method", C.c_int)]


class MxBatch(C.Structure):
    _fields_ = [("nprob", C.c_int), ("kx", _D), ("ky", _D), ("neig", C.c_int), ("tol", C.c_double),
                ("maxit", C.c_int), ("seed", C.c_ulonglong), ("E0", _D)]


class MxResult(C.Structure):
    _fields_ = [("lambda_", _D), ("relres", _D), ("absres", _D), ("iters", _I), ("converged", _I)]


class MxStats(C.Structure):
    _fields_ = [("block_steps", C.c_longlong), ("edge_vcycles", C.c_longlong),
                ("node_vcycles", C.c_longlong), ("launches", C.c_longlong),
                ("active_vectors", C.c_longlong), ("solve_seconds", C.c_double), ("device_bytes", C.c_longlong),
                ("graph_replays", C.c_longlong), ("graph_captures", C.c_longlong)]


class Grid2DStruct(C.Structure):
    _fields_ = [("hx", C.c_double), ("hy", C.c_double), ("nx", C.c_int), ("ny", C.c_int),
                ("xbloch", C.c_double), ("ybloch", C.c_double), ("xfactor", Cplx), ("yfactor", Cplx),
                ("eps", _D), ("dev", C.c_void_p)]


class Matrix(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("ld", C.c_int), ("a", C.c_void_p)]


class EpsPattern(C.Structure):
    _fields_ = [("rectangle", _D), ("rectangles", C.c_int), ("circle", _D), ("circles", C.c_int),
                ("eps_base", C.c_double)]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_VP = C.c_void_p
_GP = C.POINTER(Grid2DStruct)
_MP = C.POINTER(Matrix)
_sig("mx_solver_create", C.c_int, [C.POINTER(MxConfig), _D, C.POINTER(C.c_void_p)])
_sig("mx_solver_destroy", None, [_VP])
_sig("mx_solve", C.c_int, [_VP, C.POINTER(MxBatch), C.POINTER(MxResult)])
_sig("mx_get_vectors", C.c_int, [_VP, C.c_int, _D])
_sig("mx_eta_gradient", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, _D])
_sig("mx_extrapolate", C.c_int, [_VP, C.POINTER(MxBatch), C.c_int, C.c_int, _D, _D, _D])
_sig("mx_get_stats", C.c_int, [_VP, C.POINTER(MxStats)])
_sig("mx_time_kernel", C.c_int, [_VP, C.c_int, C.c_int, _D, _D])
# reference module API
_sig("new_grid2d", _GP, [C.c_int, C.c_int, C.c_double, C.c_double])
_sig("del_grid2d", None, [_GP])
_sig("coarsen_grid2d", _GP, [_GP])
_sig("sync_grid2d", None, [_GP])
_sig("new_edge2d", _VP, [_GP])
_sig("del_edge2d", None, [_VP])
_sig("new_node2d", _VP, [_GP])
_sig("del_node2d", None, [_VP])
_sig("denseto_edge2d", None, [_D, _VP])
_sig("densefrom_edge2d", None, [_VP, _D])
_sig("denseto_node2d", None, [_D, _VP])
_sig("densefrom_node2d", None, [_VP, _D])
_sig("zero_edge2d", None, [_VP])
_sig("copy_edge2d", None, [_VP, _VP])
_sig("scale_edge2d", None, [Cplx, _VP])
_sig("add_edge2d", None, [Cplx, _VP, _VP])
_sig("norm2_edge2d", C.c_double, [_VP])
_sig("normmax_edge2d", C.c_double, [_VP])
_sig("dotprod_edge2d", Cplx, [_VP, _VP])
_sig("center_edge2d", None, [_VP])
_sig("nullprod_edge2d", None, [_VP, C.POINTER(Cplx), C.POINTER(Cplx)])
_sig("addeval_edge2d", None, [Cplx, Cplx, _VP, _VP])
_sig("energyprod_edge2d", Cplx, [Cplx, Cplx, _VP, _VP])
_sig("massprod_edge2d", Cplx, [_VP, _VP])
_sig("gsforward_edge2d", None, [Cplx, Cplx, _VP, _VP])
_sig("gsbackward_edge2d", None, [Cplx, Cplx, _VP, _VP])
_sig("gssymm_edge2d", None, [Cplx, Cplx, _VP, _VP])
_sig("prolongation_edge2d", None, [Cplx, _VP, _VP])
_sig("restriction_edge2d", None, [Cplx, _VP, _VP])
_sig("vcycle_edge2d", None, [C.c_int, C.c_int, _MP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP)])
_sig("pinvit_edge2d", None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _VP, _MP, C.POINTER(_VP),
                             C.POINTER(_VP), C.POINTER(_VP), _MP, C.POINTER(_VP), C.POINTER(_VP),
                             C.POINTER(_VP)])
_sig("orthonormalize_edge2d", None, [C.c_int, C.POINTER(_VP), C.POINTER(Cplx), _VP])
_sig("addeval_node2d", None, [Cplx, _VP, _VP])
_sig("gsforward_node2d", None, [Cplx, _VP, _VP])
_sig("gsbackward_node2d", None, [Cplx, _VP, _VP])
_sig("prolongation_node2d", None, [Cplx, _VP, _VP])
_sig("restriction_node2d", None, [Cplx, _VP, _VP])
_sig("gradient_node2d", None, [Cplx, _VP, _VP])
_sig("adjgradient_node2d", None, [Cplx, _VP, _VP])
_sig("dotprod_node2d", Cplx, [_VP, _VP])
_sig("vcycle_node2d", None, [C.c_int, C.c_int, _MP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP)])
_sig("new_matrix", _MP, [C.c_int, C.c_int])
_sig("del_matrix", None, [_MP])
_sig("densematrix_edge2d", None, [Cplx, Cplx, _GP, _MP])
_sig("densematrix_node2d", None, [C.c_bool, _GP, _MP])
_sig("potrf_matrix", C.c_int, [_MP])
_sig("sygv_matrix", C.c_int, [_MP, _MP, _D])
_VPP = C.POINTER(_VP)
for _n in ("gsforward_simple_edge2d", "gsbackward_simple_edge2d", "gssymm_simple_edge2d"):
    _sig(_n, None, [Cplx, Cplx, _VP, _VP])
_sig("hcycle_edge2d", None, [C.c_int, C.c_int, _MP, _VPP, _VPP, _VPP, _VPP, _VPP])
for _n in ("mgcginit_edge2d", "mgcgstep_edge2d", "mgcginit_node2d", "mgcgstep_node2d"):
    _sig(_n, None, [C.c_int, C.c_int, _VP, _VP, _VP, _VP, _VPP, _VPP, _VPP])
# callbacks are passed as plain addresses (c_void_p)
_sig("interpolate_edge2d", None, [_VP, _VP, _VP])
_sig("l2functional_edge2d", None, [_VP, _VP, _VP, _VP, _VP])
_sig("l2norm_edge2d", C.c_double, [_VP, _VP, _VP])
_sig("interpolate_node2d", None, [_VP, _VP, _VP])
_sig("l2functional_node2d", None, [_VP, _VP, _VP, _VP])
_sig("l2norm_node2d", C.c_double, [_VP, _VP, _VP])
_sig("copy_node2d", None, [_VP, _VP])
_sig("zero_node2d", None, [_VP])
_PP = C.POINTER(EpsPattern)
_sig("new_epspattern", _PP, [C.c_double, C.c_int, C.c_int])
_sig("del_epspattern", None, [_PP])
_sig("setpattern_grid2d", None, [_PP, _GP])
_sig("new_ncfile", _VP, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, _PP, C.c_char_p])
_sig("del_ncfile", None, [_VP])
_sig("write_ncfile", None, [_VP, C.c_int, _D, C.POINTER(_VP)])
_sig("cairodraw_edge2d", None, [_VP, C.c_int, C.c_int, _PP, C.c_char_p])


def _p(a):
    return a.ctypes.data_as(_D)


def check(code, what="maxwell_amd"):
    if code != 0:
        raise RuntimeError("%s: error %d (%s)" % (what, code, MX_ERRORS.get(code, "?")))


class Solver:
    """Batched block PINVIT solver (mx_solver) for one permittivity field."""

    def __init__(self, eta, n0=4, levels=None, precision=MX_PREC_F64, block=12, max_problems=1,
                 smoother_steps=2, prec_steps=2, gradient_steps=5, phase_kernels=False, lock=True,
                 march_min=256, march_rows=192, march_fuse=True, autotune=False, graphs=False,
                 method=MX_METHOD_PINVIT):
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        n = eta.shape[0]
        if levels is None:
            levels = int(round(np.log2(n // n0)))
        assert (n0 << levels) == n and eta.shape == (n, n)
        self.n, self.n0, self.levels, self.block = n, n0, levels, block
        self.max_problems = max_problems
        cfg = MxConfig(n0, levels, precision, block, max_problems, smoother_steps, prec_steps,
                       gradient_steps, int(bool(phase_kernels)), int(bool(lock)), march_min, march_rows,
                       int(bool(march_fuse)), int(bool(autotune)), int(bool(graphs)), int(method))
        h = C.c_void_p()
        check(lib.mx_solver_create(C.byref(cfg), _p(eta), C.byref(h)), "mx_solver_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.mx_solver_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def nedge(self):
        return 2 * self.n * self.n

    def solve(self, kx, ky, neig, tol=1e-8, maxit=200, seed=1, E0=None):
        kx = np.ascontiguousarray(np.atleast_1d(kx), dtype=np.float64)
        ky = np.ascontiguousarray(np.atleast_1d(ky), dtype=np.float64)
        P, m = len(kx), self.block
        lam, rel, ab = np.zeros(P * m), np.zeros(P * m), np.zeros(P * m)
        its, conv = np.zeros(P, np.int32), np.zeros(P, np.int32)
        e0 = None
        if E0 is not None:
            e0 = np.ascontiguousarray(E0, dtype=np.complex128).reshape(-1)
            assert e0.size == P * m * self.nedge
        b = MxBatch(P, _p(kx), _p(ky), neig, tol, maxit, seed,
                    e0.view(np.float64).ctypes.data_as(_D) if e0 is not None else None)
        r = MxResult(_p(lam), _p(rel), _p(ab), its.ctypes.data_as(_I), conv.ctypes.data_as(_I))
        check(lib.mx_solve(self.h, C.byref(b), C.byref(r)), "mx_solve")
        return dict(lam=lam.reshape(P, m), relres=rel.reshape(P, m), absres=ab.reshape(P, m),
                    iters=its, converged=conv.astype(bool))

    def vectors(self, p):
        E = np.zeros(self.block * self.nedge, np.complex128)
        check(lib.mx_get_vectors(self.h, p, E.view(np.float64).ctypes.data_as(_D)), "mx_get_vectors")
        return E.reshape(self.block, self.nedge)

    def extrapolate(self, kx, ky, H, sources):
        """Rayleigh-Ritz extrapolation (exp_blochsurf.c:332-406): H has shape
        (nprob, sources * ieigs, nedge), the basis vectors of each problem
        (e.g. its current, old and very old blocks); returns the start blocks
        (nprob, block, nedge) for solve(E0=...) at (kx, ky) and their Ritz
        values of K (nprob, block)."""
        kx = np.ascontiguousarray(np.atleast_1d(kx), dtype=np.float64)
        ky = np.ascontiguousarray(np.atleast_1d(ky), dtype=np.float64)
        P, m = len(kx), self.block
        H = np.ascontiguousarray(H, dtype=np.complex128).reshape(P, -1, self.nedge)
        dim = H.shape[1]
        assert dim % sources == 0
        E0 = np.zeros((P, m, self.nedge), np.complex128)
        lam = np.zeros(P * m)
        b = MxBatch(P, _p(kx), _p(ky), m, 0.0, 0, 0, None)
        check(lib.mx_extrapolate(self.h, C.byref(b), sources, dim // sources, H.view(np.float64).ctypes.data_as(_D),
                                 E0.view(np.float64).ctypes.data_as(_D), _p(lam)), "mx_extrapolate")
        return E0, lam.reshape(P, m)

    def eta_gradient(self, p, first, count=1):
        g = np.zeros(self.n * self.n)
        check(lib.mx_eta_gradient(self.h, p, first, count, _p(g)), "mx_eta_gradient")
        return g.reshape(self.n, self.n)

    def stats(self):
        st = MxStats()
        check(lib.mx_get_stats(self.h, C.byref(st)), "mx_get_stats")
        return {k: getattr(st, k) for k, _ in MxStats._fields_}

    def tuning(self):
        """Smoother tiling per level (0 = square tiles, r = marching rows)."""
        e = (C.c_int * (self.levels + 1))()
        n = (C.c_int * (self.levels + 1))()
        check(lib.mx_get_tuning(self.h, e, n), "mx_get_tuning")
        return {"edge": list(e), "node": list(n)}

    def time_kernel(self, kind, reps):
        ms, by = C.c_double(), C.c_double()
        check(lib.mx_time_kernel(self.h, kind, reps, C.byref(ms), C.byref(by)), "mx_time_kernel")
        return ms.value, by.value


# ---------------------------------------------------------------- reference API helpers
class Grid:
    """A grid2d of the reference API with its eps set from eta (1/eps per cell)."""

    def __init__(self, eta, kx=0.0, ky=0.0, hx=None, hy=None, handle=None):
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        ny, nx = eta.shape
        if handle is None:
            handle = lib.new_grid2d(nx, ny, 1.0 / nx if hx is None else hx, 1.0 / ny if hy is None else hy)
            C.memmove(handle.contents.eps, eta.ctypes.data, eta.nbytes)
            lib.sync_grid2d(handle)
        self.g = handle
        self.nx, self.ny = nx, ny
        self.set_bloch(kx, ky)

    def set_bloch(self, kx, ky):
        s = self.g.contents
        s.xbloch, s.ybloch = kx, ky
        s.xfactor = Cplx.of(np.exp(1j * kx)) if kx != 0 else Cplx(1.0, 0.0)
        s.yfactor = Cplx.of(np.exp(1j * ky)) if ky != 0 else Cplx(1.0, 0.0)

    def coarse(self):
        h = lib.coarsen_grid2d(self.g)
        s = h.contents
        eta = np.ctypeslib.as_array(s.eps, shape=(s.ny * s.nx,)).reshape(s.ny, s.nx).copy()
        gc = Grid.__new__(Grid)
        gc.g, gc.nx, gc.ny = h, s.nx, s.ny
        return gc

    @property
    def factors(self):
        s = self.g.contents
        return s.xfactor.value(), s.yfactor.value()


class Edge:
    def __init__(self, grid, values=None):
        self.grid = grid
        self.h = C.c_void_p(lib.new_edge2d(grid.g))
        self.n = 2 * grid.nx * grid.ny
        if values is not None:
            self.set(values)

    def set(self, v):
        v = np.ascontiguousarray(v, dtype=np.complex128)
        assert v.size == self.n
        lib.denseto_edge2d(v.view(np.float64).ctypes.data_as(_D), self.h)

    def get(self):
        v = np.zeros(self.n, np.complex128)
        lib.densefrom_edge2d(self.h, v.view(np.float64).ctypes.data_as(_D))
        return v

    def __del__(self):
        if getattr(self, "h", None):
            lib.del_edge2d(self.h)
            self.h = None


class Node:
    def __init__(self, grid, values=None):
        self.grid = grid
        self.h = C.c_void_p(lib.new_node2d(grid.g))
        self.n = grid.nx * grid.ny
        if values is not None:
            self.set(values)

    def set(self, v):
        v = np.ascontiguousarray(v, dtype=np.complex128)
        assert v.size == self.n
        lib.denseto_node2d(v.view(np.float64).ctypes.data_as(_D), self.h)

    def get(self):
        v = np.zeros(self.n, np.complex128)
        lib.densefrom_node2d(self.h, v.view(np.float64).ctypes.data_as(_D))
        return v

    def __del__(self):
        if getattr(self, "h", None):
            lib.del_node2d(self.h)
            self.h = None


def header_functions():
    """Names of all functions declared in include/*.h (for export checks)."""
    import re
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(INCLUDE, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"#define[^\n]*", "", txt)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", txt):
            name = m.group(1)
            if name not in ("sizeof",):
                names.add(name)
    return sorted(names)
