"""Drop-in check: the reference's own drivers, compiled unmodified and linked
against libmaxwell_amd.so instead of the reference's grid2d/edge2d/node2d/
linalg/basic modules (recipes in oracle/Makefile, target `driver`), run on
the GPU library and print the eigenvalues of the oracle's dense pencil:
the Bloch-surface driver exp_blochsurf.c (built with its own headers) and
the Bloch-path driver exp_bloch2d.c (built with this library's headers, as
the reference's edge2d.h needs netcdf.h for its eigenvector file).

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
DRV2D = os.path.join(ROOT, "oracle", "_ref", "exp_bloch2d")
NOT_BUILT = "driver not built: needs the reference sources at build time"

pytestmark = pytest.mark.gpu


def dense_eigs(eta, kx, ky, count):
    g = O.Grid2D(eta, kx=kx, ky=ky)
    w = sl.eigh(O.dense_edge(g, 1, 0), O.dense_edge(g, 0, 1), eigvals_only=True)
    return w[np.abs(w) > 1e-6][:count]


@pytest.mark.skipif(not os.path.exists(DRV), reason=NOT_BUILT)
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


def bloch_curve(pt, steps):
    """exp_bloch2d.c:23-47: Gamma -> X -> M -> Gamma, `steps` points per edge."""
    if pt < steps:
        return np.pi * pt / steps, 0.0
    if pt < 2 * steps:
        return np.pi, np.pi * (pt - steps) / steps
    t = np.pi * (1.0 - (pt - 2 * steps) / (1.41 * steps))
    return t, t


@pytest.mark.skipif(not os.path.exists(DRV2D), reason=NOT_BUILT)
def test_reference_bloch_path_driver_on_library(tmp_path):
    """exp_bloch2d: nested iteration at Gamma, then the Bloch path with 2nd
    order Ritz extrapolation, Householder M-orthonormalisation, sign
    continuity and the composite Chern number, from an OpenMP team of 4;
    its protocol and its eigenvector file (read back with scipy) hold the
    dense pencil's eigenvalues at the path's k-points."""
    from scipy.io import netcdf_file
    n0, L, ieigs, throwaway, steps = 4, 2, 6, 2, 2
    allsteps = int(3.41 * steps + 1)
    (tmp_path / "Pictures").mkdir()
    env = dict(os.environ, maxwell_coarse=str(n0), maxwell_levels=str(L), maxwell_permittivity="c",
               maxwell_eigs=str(ieigs), maxwell_throwaway=str(throwaway), maxwell_chern="2",
               maxwell_continuity="s", maxwell_blochsteps=str(steps), maxwell_extra="2",
               maxwell_protocol=str(tmp_path / "path.dat"), maxwell_grid=str(tmp_path / "grid.png"),
               maxwell_cdf=str(tmp_path / "path.nc"), maxwell_stop="r", maxwell_residual="1e-10",
               maxwell_pinvit="150", maxwell_bloch="150", maxwell_visual="f", OMP_NUM_THREADS="4")
    print("running", DRV2D, flush=True)
    p = subprocess.run([DRV2D], env=env, cwd=str(tmp_path), stdin=subprocess.DEVNULL, capture_output=True,
                       text=True, timeout=240)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    n = n0 << L
    # the driver's circle pattern (exp_bloch2d.c:49-61): eps 4, radius 0.333, centred
    eta = O.setpattern(n, n, 1.0, circles=[(4.0, 0.5, 0.5, 0.333)])
    ref = {pt: dense_eigs(eta, *bloch_curve(pt, steps), ieigs) for pt in range(allsteps)}
    # the driver does not remove the two constant fields at Gamma (step 0):
    # they are eigenvectors of eigenvalue 0, below the first non-zero ones
    ref[0] = np.concatenate([[0.0, 0.0], ref[0][:ieigs - 2]])
    rows = [ln.split() for ln in open(tmp_path / "path.dat") if ln.strip() and not ln.startswith("#")]
    assert [int(r[0]) for r in rows] == list(range(allsteps)), rows
    for r in rows:
        lam = np.array([float(v) for v in r[1:1 + ieigs]])
        w = ref[int(r[0])]
        assert np.abs(lam - w).max() < 2e-6 + 1e-9 * w.max(), (r[0], lam, w)
    f = netcdf_file(str(tmp_path / "path.nc"), "r", mmap=False)
    try:
        assert f.dimensions["blochsteps"] == allsteps and f.dimensions["eigenvalues"] == ieigs
        assert f.dimensions["graphx"] == 64 and f.dimensions["graphy"] == 64
        assert f.dimensions["rectangles"] is None and f.dimensions["circles"] == 4
        assert np.array_equal(f.variables["circle"].data, [4.0, 0.5, 0.5, 0.333])
        lam = np.array(f.variables["lambda"].data)
        for pt in range(allsteps):
            kx, ky = bloch_curve(pt, steps)
            assert abs(f.variables["xbloch"].data[pt] - kx) < 1e-15 and abs(f.variables["ybloch"].data[pt] - ky) < 1e-15
            assert abs(f.variables["xfactorr"].data[pt] - np.cos(kx)) < 1e-15
            assert abs(f.variables["yfactori"].data[pt] - np.sin(ky)) < 1e-15
            w = ref[pt]
            assert np.abs(lam[pt] - w).max() < 1e-8 * w.max() + 1e-9, (pt, lam[pt], w)
        for name in ("vectorxr", "vectorxi", "vectoryr", "vectoryi"):
            v = np.array(f.variables[name].data)
            assert v.shape == (allsteps, ieigs, 64, 64) and np.isfinite(v).all() and np.abs(v).max() < 1e30
    finally:
        f.close()
    pic = tmp_path / "Pictures" / "eigenvector_step000_0.png"
    assert open(pic, "rb").read().startswith(b"P5\n40 40\n255\n")
