"""Independent finite-element assembly by numerical quadrature (numpy).

Builds the Bloch-periodic lowest-order Nedelec (edge) stiffness and mass
matrices and the bilinear nodal Laplacian directly from the basis functions,
with Gauss quadrature, without using any stencil coefficient.  The tests use it
to pin the oracle's stencils and dense assembly from the mathematics.

DOF conventions (the reference's, grid2d.h / edge2d.h / node2d.h): x-edge (i,j)
runs from node (i,j) to node (i+1,j) and its DOF is the line integral of the
tangential component; y-edge (i,j) runs from node (i,j) to (i,j+1).  Fields are
Bloch periodic: u(x + nx*hx, y) = xf u(x, y), u(x, y + ny*hy) = yf u(x, y).
"""
import numpy as np

_GP = np.array([-1.0, 1.0]) / np.sqrt(3.0)  # 2-point Gauss on [-1, 1]
_GW = np.array([1.0, 1.0])


def _quad_points(hx, hy):
    pts = []
    for a, wa in zip(_GP, _GW):
        for b, wb in zip(_GP, _GW):
            pts.append((0.5 * hx * (a + 1), 0.5 * hy * (b + 1), 0.25 * hx * hy * wa * wb))
    return pts


def _edge_basis(x, y, hx, hy):
    """Values (4, 2) and curls (4,) of the cell's edge basis functions at (x, y):
    order bottom x-edge, top x-edge, left y-edge, right y-edge."""
    val = np.array([
        [(1 - y / hy) / hx, 0.0],
        [(y / hy) / hx, 0.0],
        [0.0, (1 - x / hx) / hy],
        [0.0, (x / hx) / hy],
    ])
    # curl u = d u_y / dx - d u_x / dy
    cur = np.array([1.0, -1.0, -1.0, 1.0]) / (hx * hy)
    return val, cur


def _wrap(i, j, nx, ny, xf, yf):
    ph = 1.0 + 0j
    if i >= nx:
        i -= nx
        ph *= xf
    if j >= ny:
        j -= ny
        ph *= yf
    return i, j, ph


def edge_matrices(eta, hx, hy, xf, yf):
    """Dense K (with per-cell weight eta) and M for the edge space."""
    ny, nx = eta.shape
    N = nx * ny
    K = np.zeros((2 * N, 2 * N), complex)
    M = np.zeros((2 * N, 2 * N), complex)
    pts = _quad_points(hx, hy)
    ke = np.zeros((4, 4))
    me = np.zeros((4, 4))
    for (x, y, w) in pts:
        val, cur = _edge_basis(x, y, hx, hy)
        ke += w * np.outer(cur, cur)
        me += w * val @ val.T
    for j in range(ny):
        for i in range(nx):
            loc = []
            for (ii, jj, off) in ((i, j, 0), (i, j + 1, 0), (i, j, N), (i + 1, j, N)):
                a, b, ph = _wrap(ii, jj, nx, ny, xf, yf)
                loc.append((off + a + b * nx, ph))
            for p, (gp, php) in enumerate(loc):
                for q, (gq, phq) in enumerate(loc):
                    K[gp, gq] += np.conj(php) * eta[j, i] * ke[p, q] * phq
                    M[gp, gq] += np.conj(php) * me[p, q] * phq
    return K, M


def node_laplacian(nx, ny, hx, hy, xf, yf):
    """Dense bilinear (Q1) Laplacian on the nodes."""
    N = nx * ny
    A = np.zeros((N, N), complex)
    ke = np.zeros((4, 4))
    for (x, y, w) in _quad_points(hx, hy):
        # local nodes (0,0), (1,0), (0,1), (1,1)
        gx = np.array([-(1 - y / hy), (1 - y / hy), -(y / hy), (y / hy)]) / hx
        gy = np.array([-(1 - x / hx), -(x / hx), (1 - x / hx), (x / hx)]) / hy
        ke += w * (np.outer(gx, gx) + np.outer(gy, gy))
    for j in range(ny):
        for i in range(nx):
            loc = []
            for p in range(4):
                a, b, ph = _wrap(i + (p & 1), j + (p >> 1), nx, ny, xf, yf)
                loc.append((a + b * nx, ph))
            for p, (gp, php) in enumerate(loc):
                for q, (gq, phq) in enumerate(loc):
                    A[gp, gq] += np.conj(php) * ke[p, q] * phq
    return A


def gradient_matrix(nx, ny, xf, yf):
    """G: node values -> edge line integrals of their gradient."""
    N = nx * ny
    G = np.zeros((2 * N, N), complex)
    for j in range(ny):
        for i in range(nx):
            a, b, ph = _wrap(i + 1, j, nx, ny, xf, yf)
            G[i + j * nx, a + b * nx] += ph
            G[i + j * nx, i + j * nx] -= 1
            a, b, ph = _wrap(i, j + 1, nx, ny, xf, yf)
            G[N + i + j * nx, a + b * nx] += ph
            G[N + i + j * nx, i + j * nx] -= 1
    return G


def empty_lattice_eigs(nx, ny, hx, hy, kx, ky, eta=1.0):
    """Closed-form non-zero eigenvalues of the uniform-eta pencil (K, M)."""
    lam = []
    for p in range(nx):
        tx = (2 * np.pi * p + kx) / nx
        for q in range(ny):
            ty = (2 * np.pi * q + ky) / ny
            lam.append(eta * (12 / hx ** 2 * np.sin(tx / 2) ** 2 / (2 + np.cos(tx))
                              + 12 / hy ** 2 * np.sin(ty / 2) ** 2 / (2 + np.cos(ty))))
    return np.sort(np.array(lam))
