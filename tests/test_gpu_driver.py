"""Drop-in check: the reference's own Bloch-surface driver (exp_blochsurf.c
and parameters.c, compiled unmodified with their own headers and linked
against libmaxwell_amd.so instead of the reference's grid2d/edge2d/node2d/
linalg/basic modules; recipe in oracle/Makefile, target `driver`) runs on
the GPU library and prints the eigenvalues of the oracle's dense pencil.

The driver reads its parameters from maxwell_* environment variables
(parameters.c), writes its protocol file (one line per Bloch point:
x index, y index, lambda - 1 for every block vector, steps) and uses the
module API from an OpenMP team of 4 threads."""
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "oracle", "_ref", "exp_blochsurf")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRV),
                                 reason="driver not built: needs the reference sources at build time")]


def test_reference_driver_on_library(tmp_path):
    n0, L, ieigs, throwaway, steps = 4, 2, 6, 2, 4
    env = dict(os.environ, maxwell_coarse=str(n0), maxwell_levels=str(L), maxwell_permittivity="c",
               maxwell_eigs=str(ieigs), maxwell_throwaway=str(throwaway), maxwell_blochsteps=str(steps),
               maxwell_protocol=str(tmp_path / "surf.dat"), maxwell_grid=str(tmp_path / "grid.png"),
               maxwell_pinvit="150", maxwell_bloch="150", maxwell_residual="1e-10", OMP_NUM_THREADS="4")
    print("running", DRV, flush=True)
    p = subprocess.run([DRV], env=env, cwd=str(tmp_path), stdin=subprocess.DEVNULL, capture_output=True,
                       text=True, timeout=240)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    rows = [ln.split() for ln in open(tmp_path / "surf.dat") if ln.strip() and not ln.startswith("#")]
    assert len(rows) == steps * steps, rows
    n = n0 << L
    # the driver's circle pattern: eps 4 in the circle of radius 0.333 at the centre
    eta = O.setpattern(n, n, 1.0, circles=[(4.0, 0.5, 0.5, 0.333)])
    seen = set()
    for r in rows:
        ix, iy = int(r[0]), int(r[1])
        seen.add((ix, iy))
        lam = np.array([float(v) for v in r[2:2 + ieigs + throwaway]])
        kx = 2 * np.pi * ix / (steps - 1) - np.pi
        ky = 2 * np.pi * iy / (steps - 1) - np.pi
        g = O.Grid2D(eta, kx=kx, ky=ky)
        w = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1), eigvals_only=True)
        w = w[np.abs(w) > 1e-6][:ieigs]
        # printed with %f: six decimals
        assert np.abs(lam[:ieigs] - w).max() < 2e-6 + 1e-9 * w.max(), (ix, iy, lam[:ieigs], w)
    assert len(seen) == steps * steps
