"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the CPU-baseline leg of ``bench.py``.  The product package never imports it.

Arrays are numpy complex128 in the compact layout described in
``maxwell_oracle.h``: an edge vector is ``[x-edges (ny, nx) | y-edges (ny, nx)]``
flattened, a node vector is ``(ny, nx)`` flattened, eta is ``(ny, nx)``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "liboracle.so"])


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


_lib = _load()

_D = C.POINTER(C.c_double)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("hx", C.c_double), ("hy", C.c_double),
                ("xf", C.c_double * 2), ("yf", C.c_double * 2), ("eta", _D)]


def _p(a):
    return a.ctypes.data_as(_D)


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_G = C.POINTER(Grid)
_sig("og_setpattern", None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                             C.c_int, _D, C.c_int, _D, _D])
_sig("og_coarsen", None, [C.c_int, C.c_int, _D, _D])
_sig("og_addeval", None, [_G, C.c_double, C.c_double, C.c_double, C.c_double, _D, _D, C.c_int])
_sig("og_energyprod", None, [_G, C.c_double, C.c_double, C.c_double, C.c_double, _D, _D, C.c_int, _D])
_sig("og_massprod", None, [_G, _D, _D, _D])
_sig("og_center", None, [_G, _D])
_sig("og_patch_solve", None, [_G, C.c_double, C.c_double, _D, _D, C.c_int, C.c_int])
_sig("og_gs_edge", None, [_G, C.c_double, C.c_double, _D, _D, C.c_int, C.c_int])
_sig("og_prolong_edge", None, [_G, C.c_double, C.c_double, _D, _D])
_sig("og_restrict_edge", None, [_G, C.c_double, C.c_double, _D, _D])
_sig("og_dense_edge", None, [_G, C.c_double, C.c_double, C.c_double, C.c_double, _D])
_sig("og_dense_node", None, [_G, C.c_int, _D])
_sig("og_potrf", C.c_int, [C.c_int, _D])
_sig("og_potrs", None, [C.c_int, _D, _D])
_sig("og_hegv", C.c_int, [C.c_int, _D, _D, _D])
_sig("og_addeval_node", None, [_G, C.c_double, C.c_double, _D, _D])
_sig("og_gs_node", None, [_G, C.c_double, _D, _D, C.c_int, C.c_int])
_sig("og_prolong_node", None, [_G, C.c_double, C.c_double, _D, _D])
_sig("og_restrict_node", None, [_G, C.c_double, C.c_double, _D, _D])
_sig("og_gradient", None, [_G, C.c_double, C.c_double, _D, _D])
_sig("og_adjgradient", None, [_G, C.c_double, C.c_double, _D, _D])
_sig("og_mg_new", C.c_void_p, [C.c_int, C.c_int, _D, C.c_int, C.c_int])
_sig("og_mg_free", None, [C.c_void_p])
_sig("og_mg_set_bloch", None, [C.c_void_p, C.c_double, C.c_double])
_sig("og_mg_vcycle_edge", None, [C.c_void_p, C.c_int, _D, _D])
_sig("og_mg_vcycle_node", None, [C.c_void_p, C.c_int, _D, _D])
_sig("og_mg_eta", _D, [C.c_void_p, C.c_int])
_sig("og_pinvit_vector", None, [C.c_void_p, C.c_int, C.c_int, C.c_double, _D])
_sig("og_lift", None, [C.c_void_p, C.c_int, _D])
_sig("og_block_solve", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                  C.c_int, C.c_int, _D, _D, _D, _D])
_sig("og_eta_gradient", None, [_G, _D, _D])


class Grid2D:
    """An nx x ny periodic grid on [0, nx*hx) x [0, ny*hy) with Bloch factors."""

    def __init__(self, eta, hx=None, hy=None, kx=0.0, ky=0.0):
        self.eta = np.ascontiguousarray(eta, dtype=np.float64)
        self.ny, self.nx = self.eta.shape
        self.hx = 1.0 / self.nx if hx is None else hx
        self.hy = 1.0 / self.ny if hy is None else hy
        self.set_bloch(kx, ky)

    def set_bloch(self, kx, ky):
        self.kx, self.ky = kx, ky
        self.xf = complex(np.cos(kx), np.sin(kx))
        self.yf = complex(np.cos(ky), np.sin(ky))

    def set_factors(self, xf, yf):
        self.xf, self.yf = complex(xf), complex(yf)

    @property
    def c(self):
        g = Grid()
        g.nx, g.ny, g.hx, g.hy = self.nx, self.ny, self.hx, self.hy
        g.xf[0], g.xf[1] = self.xf.real, self.xf.imag
        g.yf[0], g.yf[1] = self.yf.real, self.yf.imag
        g.eta = _p(self.eta)
        self._keep = g
        return C.byref(g)

    @property
    def nedge(self):
        return 2 * self.nx * self.ny

    @property
    def nnode(self):
        return self.nx * self.ny

    def coarse(self):
        ec = np.empty((self.ny // 2, self.nx // 2))
        _lib.og_coarsen(self.nx, self.ny, _p(self.eta), _p(ec))
        g = Grid2D(ec, 2 * self.hx, 2 * self.hy)
        g.set_factors(self.xf, self.yf)
        return g


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def setpattern(nx, ny, eps_base=1.0, rects=(), circles=(), hx=None, hy=None):
    hx = 1.0 / nx if hx is None else hx
    hy = 1.0 / ny if hy is None else hy
    r = np.ascontiguousarray(np.array(rects, dtype=np.float64).reshape(-1))
    c = np.ascontiguousarray(np.array(circles, dtype=np.float64).reshape(-1))
    r = r if r.size else np.zeros(1)
    c = c if c.size else np.zeros(1)
    eta = np.empty((ny, nx))
    _lib.og_setpattern(nx, ny, hx, hy, eps_base, len(rects), _p(r), len(circles), _p(c), _p(eta))
    return eta


def addeval(g, alpha, beta, x, y=None, quirk=False):
    x = _c(x)
    y = np.zeros_like(x) if y is None else _c(y).copy()
    a, b = complex(alpha), complex(beta)
    _lib.og_addeval(g.c, a.real, a.imag, b.real, b.imag, _p(x), _p(y), int(quirk))
    return y


def energyprod(g, alpha, beta, x, y, quirk=False):
    x, y = _c(x), _c(y)
    out = np.zeros(2)
    a, b = complex(alpha), complex(beta)
    _lib.og_energyprod(g.c, a.real, a.imag, b.real, b.imag, _p(x), _p(y), int(quirk), _p(out))
    return complex(out[0], out[1])


def massprod(g, x, y):
    x, y = _c(x), _c(y)
    out = np.zeros(2)
    _lib.og_massprod(g.c, _p(x), _p(y), _p(out))
    return complex(out[0], out[1])


def center(g, x):
    x = _c(x).copy()
    _lib.og_center(g.c, _p(x))
    return x


def patch_solve(g, b, x, i, j, alpha=1.0, beta=1.0):
    b = _c(b)
    x = _c(x).copy()
    _lib.og_patch_solve(g.c, alpha, beta, _p(b), _p(x), i, j)
    return x


def gs_edge(g, b, x, backward=False, colour_order=1, alpha=1.0, beta=1.0):
    b = _c(b)
    x = _c(x).copy()
    _lib.og_gs_edge(g.c, alpha, beta, _p(b), _p(x), int(backward), int(colour_order))
    return x


def prolong_edge(gf, c, f=None, alpha=1.0):
    c = _c(c)
    f = np.zeros(gf.nedge, complex) if f is None else _c(f).copy()
    a = complex(alpha)
    _lib.og_prolong_edge(gf.c, a.real, a.imag, _p(c), _p(f))
    return f


def restrict_edge(gf, f, c=None, alpha=1.0):
    f = _c(f)
    c = np.zeros(gf.nedge // 4, complex) if c is None else _c(c).copy()
    a = complex(alpha)
    _lib.og_restrict_edge(gf.c, a.real, a.imag, _p(f), _p(c))
    return c


def dense_edge(g, alpha=1.0, beta=1.0):
    """Column-major dense (alpha K + beta M), returned as a numpy (n, n) matrix A
    with (A @ x) == addeval(alpha, beta, x)."""
    n = g.nedge
    a = np.zeros(n * n, complex)
    al, be = complex(alpha), complex(beta)
    _lib.og_dense_edge(g.c, al.real, al.imag, be.real, be.imag, _p(a))
    return a.reshape(n, n).T.copy()


def dense_node(g, fixnull=True):
    n = g.nnode
    a = np.zeros(n * n, complex)
    _lib.og_dense_node(g.c, int(fixnull), _p(a))
    return a.reshape(n, n).T.copy()


def potrf(a):
    """Lower Cholesky factor of the Hermitian matrix a (numpy (n, n))."""
    n = a.shape[0]
    buf = np.ascontiguousarray(np.asarray(a, complex).T).reshape(-1).copy()
    info = _lib.og_potrf(n, _p(buf))
    return buf.reshape(n, n).T.copy(), info


def potrs(l, b):
    n = l.shape[0]
    lb = np.ascontiguousarray(np.asarray(l, complex).T).reshape(-1).copy()
    x = _c(b).copy()
    _lib.og_potrs(n, _p(lb), _p(x))
    return x


def hegv(a, b):
    """Generalised Hermitian eigenproblem from the lower triangles of a and b.
    Returns (w ascending, X) with X^H b X = I."""
    n = a.shape[0]
    ab = np.ascontiguousarray(np.asarray(a, complex).T).reshape(-1).copy()
    bb = np.ascontiguousarray(np.asarray(b, complex).T).reshape(-1).copy()
    w = np.zeros(n)
    info = _lib.og_hegv(n, _p(ab), _p(bb), _p(w))
    if info:
        raise np.linalg.LinAlgError("og_hegv info=%d" % info)
    return w, ab.reshape(n, n).T.copy()


def addeval_node(g, alpha, x, y=None):
    x = _c(x)
    y = np.zeros_like(x) if y is None else _c(y).copy()
    a = complex(alpha)
    _lib.og_addeval_node(g.c, a.real, a.imag, _p(x), _p(y))
    return y


def gs_node(g, b, x, backward=False, colour_order=1, alpha=1.0):
    b = _c(b)
    x = _c(x).copy()
    _lib.og_gs_node(g.c, alpha, _p(b), _p(x), int(backward), int(colour_order))
    return x


def prolong_node(gf, c, f=None, alpha=1.0):
    c = _c(c)
    f = np.zeros(gf.nnode, complex) if f is None else _c(f).copy()
    a = complex(alpha)
    _lib.og_prolong_node(gf.c, a.real, a.imag, _p(c), _p(f))
    return f


def restrict_node(gf, f, c=None, alpha=1.0):
    f = _c(f)
    c = np.zeros(gf.nnode // 4, complex) if c is None else _c(c).copy()
    a = complex(alpha)
    _lib.og_restrict_node(gf.c, a.real, a.imag, _p(f), _p(c))
    return c


def gradient(g, v, e=None, alpha=1.0):
    v = _c(v)
    e = np.zeros(g.nedge, complex) if e is None else _c(e).copy()
    a = complex(alpha)
    _lib.og_gradient(g.c, a.real, a.imag, _p(v), _p(e))
    return e


def adjgradient(g, e, v=None, alpha=1.0):
    e = _c(e)
    v = np.zeros(g.nnode, complex) if v is None else _c(v).copy()
    a = complex(alpha)
    _lib.og_adjgradient(g.c, a.real, a.imag, _p(e), _p(v))
    return v


def eta_gradient(g, e):
    e = _c(e)
    out = np.zeros((g.ny, g.nx))
    _lib.og_eta_gradient(g.c, _p(e), _p(out))
    return out


class Multigrid:
    """Edge and node multigrid hierarchy with n0 x n0 coarse grid and L
    refinements (finest grid n0*2^L), plus the block PINVIT solver."""

    def __init__(self, n0, L, eta_fine, colour_order=1, quirk=False):
        self.n0, self.L = n0, L
        self.n = n0 << L
        self.eta = np.ascontiguousarray(eta_fine, dtype=np.float64)
        assert self.eta.shape == (self.n, self.n)
        self.h = C.c_void_p(_lib.og_mg_new(n0, L, _p(self.eta), int(colour_order), int(quirk)))
        if not self.h:
            raise ValueError("too many levels")
        self.kx = self.ky = 0.0

    def __del__(self):
        if getattr(self, "h", None):
            _lib.og_mg_free(self.h)
            self.h = None

    def set_bloch(self, kx, ky):
        self.kx, self.ky = float(kx), float(ky)
        _lib.og_mg_set_bloch(self.h, self.kx, self.ky)

    def grid(self, l=None):
        l = self.L if l is None else l
        n = self.n0 << l
        eta = np.ctypeslib.as_array(_lib.og_mg_eta(self.h, l), shape=(n * n,)).reshape(n, n).copy()
        return Grid2D(eta, kx=self.kx, ky=self.ky)

    @property
    def nedge(self):
        return 2 * self.n * self.n

    def vcycle_edge(self, b, x=None, nu=2):
        b = _c(b)
        x = np.zeros_like(b) if x is None else _c(x).copy()
        _lib.og_mg_vcycle_edge(self.h, nu, _p(b), _p(x))
        return x

    def vcycle_node(self, b, x=None, nu=2):
        b = _c(b)
        x = np.zeros_like(b) if x is None else _c(x).copy()
        _lib.og_mg_vcycle_node(self.h, nu, _p(b), _p(x))
        return x

    def pinvit_vector(self, lam, e, prec_steps=2, gradient_steps=5):
        e = _c(e).copy()
        _lib.og_pinvit_vector(self.h, prec_steps, gradient_steps, lam, _p(e))
        return e

    def lift(self, e, gradient_steps=5):
        e = _c(e).copy()
        _lib.og_lift(self.h, gradient_steps, _p(e))
        return e

    def block_solve(self, E, neig, tol=1e-8, maxit=200, prec_steps=2, gradient_steps=5):
        """E: (blk, nedge) start block.  Returns (lam_K, E, relres, absres, steps)."""
        E = np.ascontiguousarray(E, dtype=np.complex128).copy()
        blk = E.shape[0]
        lam = np.zeros(blk)
        rel = np.zeros(blk)
        ab = np.zeros(blk)
        it = _lib.og_block_solve(self.h, neig, blk, tol, maxit, prec_steps, gradient_steps,
                                 _p(E), _p(lam), _p(rel), _p(ab))
        return lam, E, rel, ab, it


def unit_indices(idx, nx, ny):
    """x-edge (i, j) of the idx-th unit vector: odd rows j = 1, 3, ..., taken
    row by row (edge2d.c:6154-6168)."""
    i, j = idx, 1
    while j < ny and i >= nx:
        i -= nx
        j += 2
    assert j < ny
    return i, j


def orthonormalize(g, X):
    """Householder M-orthonormalisation of the rows of X (edge2d.c:6154-6296):
    unit vectors on x-edges of odd rows scaled to unit M-norm; step i builds
    the Householder vector from X[i] (alpha = |x|_M with the sign of
    Re(unit^H M x), x += alpha unit, tau = 1 / (alpha^2 + beta alpha)), applies
    it to the later vectors and removes their unit-i component; then output i
    is H_0 ... H_i unit_i.  Returns (outputs, tau, Householder vectors).  The
    unit's row above the last one is read through the Bloch factor."""
    X = np.array(X, dtype=np.complex128)
    k = len(X)
    nx, ny = g.nx, g.ny
    mx = g.hy / g.hx / 6.0
    sc = np.sqrt(0.25 / mx)

    def unit_prod(i, j, y):  # unit(i, j)^H M y
        up = y[i + (j + 1) * nx] if j + 1 < ny else g.yf * y[i]
        return np.sqrt(0.25 * mx) * (4.0 * y[i + j * nx] + y[i + (j - 1) * nx] + up)

    tau = np.zeros(k, np.complex128)
    for i in range(k):
        ui, uj = unit_indices(i, nx, ny)
        alpha = np.sqrt(massprod(g, X[i], X[i]).real)
        if alpha == 0.0:
            tau[i] = 0.0
        else:
            beta = unit_prod(ui, uj, X[i])
            if beta.real < 0.0:
                alpha = -alpha
            X[i, ui + uj * nx] += alpha * sc
            tau[i] = 1.0 / (alpha * alpha + beta * alpha)
        for j in range(i + 1, k):
            X[j] -= tau[i] * massprod(g, X[i], X[j]) * X[i]
            X[j, ui + uj * nx] -= unit_prod(ui, uj, X[j]) * sc
    out = np.zeros_like(X)
    for i in range(k - 1, -1, -1):
        ui, uj = unit_indices(i, nx, ny)
        v = np.zeros(X.shape[1], np.complex128)
        v[ui + uj * nx] = sc
        for j in range(i, -1, -1):
            v -= np.conj(tau[j]) * massprod(g, X[j], v) * X[j]
        out[i] = v
    return out, tau, X


def rayleigh_extrapolation(g, basis, eigs):
    """Rayleigh-Ritz extrapolation (exp_blochsurf.c:332-406; exp_bloch2d.c:
    825-894): the rows of `basis` (sources x ieigs vectors, e.g. the current,
    old and very old eigenvector blocks) are M-orthonormalised by Householder
    reflections at g's Bloch parameters (orthonormalize above), the Ritz
    matrix Ar[i, j] = energyprod(1, 1, w_i, w_j) of K + M is formed over the
    orthonormal w (lower triangle, as the reference fills it), its Hermitian
    eigenproblem is solved (syev_matrix, linalg.c:254-317; numpy's eigh here)
    and e_j = sum_i Ar[i, j] w_i for the lowest eigs.  Returns (e, Ritz
    values of K + M)."""
    w, _, _ = orthonormalize(g, basis)
    d = len(w)
    A = np.zeros((d, d), np.complex128)
    for j in range(d):
        for i in range(j, d):
            A[i, j] = energyprod(g, 1.0, 1.0, w[i], w[j])
    lam, C_ = np.linalg.eigh(A, UPLO="L")
    return C_[:, :eigs].T @ w, lam[:eigs]


def rod_eta(n, eps=8.9, r=0.2):
    """1/eps of a centred circular rod, sampled at cell corners (grid2d.c:195-197)."""
    return setpattern(n, n, 1.0, circles=[(eps, 0.5, 0.5, r)])


def random_block(blk, nedge, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((blk, nedge)) + 1j * rng.standard_normal((blk, nedge))).astype(np.complex128)


def graph_sample(nx, ny, hx, hy, xf, yf, v, gx, gy):
    """The field of edge vector v (dense order) at the centres of a gx x gy
    graph grid, as write_ncfile stores it (edge2d.c:582-611): for sample
    (ig, jg), i = int((ig + 0.5) nx / gx), xoff = (ig + 0.5) nx hx / gx - i hx,
    vx = (1 - xoff) x(i, j) + xoff x(i+1, j), and the same in y.  The
    neighbour past the last column (row) is the Bloch continuation
    xf x(0, j) (yf y(i, 0)).  Returns (vx, vy), each gx x gy."""
    xe = v[:nx * ny].reshape(ny, nx)
    ye = v[nx * ny:].reshape(ny, nx)
    hgx, hgy = (nx * hx) / gx, (ny * hy) / gy
    vx = np.zeros((gx, gy), complex)
    vy = np.zeros((gx, gy), complex)
    for jg in range(gy):
        j = int((jg + 0.5) * ny / gy)
        yoff = (jg + 0.5) * hgy - j * hy
        for ig in range(gx):
            i = int((ig + 0.5) * nx / gx)
            xoff = (ig + 0.5) * hgx - i * hx
            x1 = xe[j, i + 1] if i + 1 < nx else xf * xe[j, 0]
            y1 = ye[j + 1, i] if j + 1 < ny else yf * ye[0, i]
            vx[ig, jg] = (1.0 - xoff) * xe[j, i] + xoff * x1
            vy[ig, jg] = (1.0 - yoff) * ye[j, i] + yoff * y1
    return vx, vy


def gs_simple_dense(A, b, x, order):
    """Pointwise Gauss-Seidel on a dense matrix A (the operator of the edge
    space, dense order) in the given order of unknowns: x_e += (b - A x)_e /
    A_ee, one unknown at a time (edge2d.c:3892-5600 is this with its row-by-row
    order)."""
    x = np.array(x, dtype=complex)
    for e in order:
        x[e] += (b[e] - A[e] @ x) / A[e, e]
    return x


def simple_colour_order(nx, ny, backward=False):
    """The four-colour order of the GPU smoother: x-edges of even rows, of odd
    rows, y-edges of even columns, of odd columns (dense indices)."""
    N = nx * ny
    cols = [[i + j * nx for j in range(c, ny, 2) for i in range(nx)] for c in (0, 1)]
    cols += [[N + i + j * nx for j in range(ny) for i in range(c, nx, 2)] for c in (0, 1)]
    if backward:
        cols = cols[::-1]
    return [e for c in cols for e in c]


def _edge_mid_values(func, nx, ny, hx, hy, xrows, ycols):
    """func at x-edge midpoints ((i+1/2) hx, j hy), j < xrows, and at y-edge
    midpoints (i hx, (j+1/2) hy), i < ycols: arrays [j, i, component]."""
    fx = np.array([[func(((i + 0.5) * hx, j * hy)) for i in range(nx)] for j in range(xrows)], dtype=complex)
    fy = np.array([[func((i * hx, (j + 0.5) * hy)) for i in range(ycols)] for j in range(ny)], dtype=complex)
    return fx, fy


def interpolate_edge(func, nx, ny, hx, hy):
    """edge2d.c:643-680: h times the tangential component at edge midpoints."""
    fx, fy = _edge_mid_values(func, nx, ny, hx, hy, ny, nx)
    return np.concatenate([(hx * fx[:, :, 0]).ravel(), (hy * fy[:, :, 1]).ravel()])


def l2functional_edge(func, nx, ny, hx, hy, xf, yf):
    """edge2d.c:683-805 in array form: an x-edge gets hy (1/2 f_x at its
    midpoint + 1/8 f_x at the midpoints of the four y-edges through its end
    nodes); across the seam the values carry xf (right) and conj(yf) (below);
    y-edges alike with conj(xf) (left) and yf (above)."""
    fx, fy = _edge_mid_values(func, nx, ny, hx, hy, ny, nx)
    fyx = fy[:, :, 0]  # x component at y-edge midpoints [j, i]
    right = np.roll(fyx, -1, axis=1).copy()
    right[:, -1] *= xf
    below = np.roll(fyx, 1, axis=0).copy()
    below[0, :] *= np.conj(yf)
    below_right = np.roll(right, 1, axis=0).copy()
    below_right[0, :] *= np.conj(yf)
    gx = hy * (0.5 * fx[:, :, 0] + 0.125 * (fyx + right + below + below_right))
    fxy = fx[:, :, 1]  # y component at x-edge midpoints
    left = np.roll(fxy, 1, axis=1).copy()
    left[:, 0] *= np.conj(xf)
    above = np.roll(fxy, -1, axis=0).copy()
    above[-1, :] *= yf
    above_left = np.roll(left, -1, axis=0).copy()
    above_left[-1, :] *= yf
    gy = hx * (0.5 * fy[:, :, 1] + 0.125 * (fxy + left + above + above_left))
    return np.concatenate([gx.ravel(), gy.ravel()])


def l2norm_edge(v, func, nx, ny, hx, hy, xf, yf):
    """edge2d.c:923-1206 in array form: per cell, the four edge midpoints with
    weight hx hy / 4 each, |f - u_h|^2 with u_h the cell's edge element
    field (tangential component from the edge itself, normal component the
    mean of the two opposite edges of the cell).  Seam values as the
    reference takes them: the top line's x-edges are conj(yf) x(i, 0), the
    right column's y-edges xf y(0, j) except conj(xf) y(0, ny-1)."""
    X = v[:nx * ny].reshape(ny, nx)
    Y = v[nx * ny:].reshape(ny, nx)
    f0 = (lambda p: (0.0, 0.0)) if func is None else func
    fx, fy = _edge_mid_values(f0, nx, ny, hx, hy, ny + 1, nx + 1)
    s = 0.0
    Xe = np.vstack([X, np.conj(yf) * X[:1]])
    wx = np.full(ny + 1, 0.5)
    wx[0] = wx[-1] = 0.25
    s += (wx[:, None] * np.abs(fx[:, :, 0] - Xe / hx) ** 2).sum()
    Yr = np.roll(Y, -1, axis=1).copy()
    Yr[:, -1] = xf * Y[:, 0]
    my = (Y + Yr) / (2 * hy)  # cell (i, j)
    s += 0.25 * (np.abs(fx[:-1, :, 1] - my) ** 2).sum() + 0.25 * (np.abs(fx[1:, :, 1] - my) ** 2).sum()
    last = np.full(ny, xf, dtype=complex)
    last[-1] = np.conj(xf)
    Ye = np.hstack([Y, (last * Y[:, 0])[:, None]])
    wy = np.full(nx + 1, 0.5)
    wy[0] = wy[-1] = 0.25
    s += (wy[None, :] * np.abs(fy[:, :, 1] - Ye / hy) ** 2).sum()
    Xt = np.roll(X, -1, axis=0).copy()
    Xt[-1, :] = yf * X[0, :]
    mx = (X + Xt) / (2 * hx)
    s += 0.25 * (np.abs(fy[:, :-1, 0] - mx) ** 2).sum() + 0.25 * (np.abs(fy[:, 1:, 0] - mx) ** 2).sum()
    return np.sqrt(s * hx * hy)


_GQ = (0.5 * (1 - np.sqrt(1 / 3)), 0.5 * (1 + np.sqrt(1 / 3)))


def l2functional_node(func, nx, ny, hx, hy, xf, yf):
    """node2d.c:118-303: 2 x 2 Gauss rule per cell, bilinear basis, weight
    hx hy / 4 per point; a cell across the seam contributes to the wrapped
    node with conj(xf), conj(yf)."""
    out = np.zeros((ny, nx), complex)
    for t in _GQ:
        for s in _GQ:
            fq = np.array([[func(((i + s) * hx, (j + t) * hy)) for i in range(nx)] for j in range(ny)], complex)
            for b in (0, 1):
                for a in (0, 1):
                    w = (s if a else 1 - s) * (t if b else 1 - t)
                    c = w * fq
                    if a:
                        c = np.roll(c, 1, axis=1)
                        c[:, 0] *= np.conj(xf)
                    if b:
                        c = np.roll(c, 1, axis=0)
                        c[0, :] *= np.conj(yf)
                    out += c
    return (0.25 * hx * hy * out).ravel()


def l2norm_node(v, func, nx, ny, hx, hy, xf, yf):
    """node2d.c:305-385: 2 x 2 Gauss rule per cell of |u_h - f|^2, u_h the
    bilinear interpolant of the cell's corner values (xf, yf across the
    seam)."""
    V = v.reshape(ny, nx)
    R = np.roll(V, -1, axis=1).copy()
    R[:, -1] *= xf
    T = np.roll(V, -1, axis=0).copy()
    T[-1, :] *= yf
    RT = np.roll(R, -1, axis=0).copy()
    RT[-1, :] *= yf
    s2 = 0.0
    for t in _GQ:
        for s in _GQ:
            u = (1 - s) * (1 - t) * V + s * (1 - t) * R + (1 - s) * t * T + s * t * RT
            if func is not None:
                u = u - np.array([[func(((i + s) * hx, (j + t) * hy)) for i in range(nx)] for j in range(ny)])
            s2 += (np.abs(u) ** 2).sum()
    return np.sqrt(s2 * 0.25 * hx * hy)


def field_edge(p):
    """The test field og_field_edge (the same formula, numpy)."""
    x, y = p
    tau = 6.283185307179586
    return (np.cos(tau * y) + 0.2j * x + 0.1 * x * y, np.sin(tau * x) * (1.0 + 0.5j) + y)


def field_node(p):
    """The test field og_field_node (the same formula, numpy)."""
    x, y = p
    tau = 6.283185307179586
    return (1.0 + 0.5 * x) * np.exp(1j * tau * (x + 2.0 * y)) + 0.3j * y * y


def field_pointer(name):
    """Address of one of the oracle's C test fields, to pass as a callback."""
    return C.cast(getattr(_lib, name), C.c_void_p).value
