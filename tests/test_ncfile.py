"""The drivers' eigenvector file (new_ncfile / write_ncfile / del_ncfile,
edge2d.c:385-640), written without libnetcdf, read back with scipy's
independent netCDF reader: the reference's dimensions and variables, the
fill value in what is not written yet, the pattern, and -- on the GPU, from
device vectors -- the Bloch data, lambda - 1 and the sampled fields, which
must equal the oracle's restatement of the sampling."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest
from scipy.io import netcdf_file

import maxwell_amd as M
import oracle as O

FILL = 9.9692099683868690e+36
VARS = ["xbloch", "ybloch", "xfactorr", "xfactori", "yfactorr", "yfactori", "rectangle", "circle", "lambda",
        "vectorxr", "vectorxi", "vectoryr", "vectoryi"]


def pattern(rects, circles):
    p = M.lib.new_epspattern(1.0, len(rects), len(circles))
    for k, r in enumerate(rects):
        for q in range(5):
            p.contents.rectangle[5 * k + q] = r[q]
    for k, c in enumerate(circles):
        for q in range(4):
            p.contents.circle[4 * k + q] = c[q]
    return p


def read(path):
    f = netcdf_file(str(path), "r", mmap=False)
    out = dict(version=f.version_byte, dims=dict(f.dimensions),
               vars={k: (v.dimensions, np.array(v.data)) for k, v in f.variables.items()})
    f.close()
    return out


def test_header_dimensions_and_fill(tmp_path):
    rects = [(2.0, 0.1, 0.2, 0.3, 0.4)]
    circles = [(8.9, 0.5, 0.5, 0.2), (4.0, 0.25, 0.75, 0.1)]
    pat = pattern(rects, circles)
    path = tmp_path / "a.nc"
    nf = M.lib.new_ncfile(None, 3, 5, 8, 6, pat, str(path).encode())
    M.lib.del_ncfile(nf)
    M.lib.del_epspattern(pat)
    r = read(path)
    assert r["version"] == 2
    assert r["dims"] == dict(blochsteps=5, eigenvalues=3, graphx=8, graphy=6, rectangles=5, circles=8)
    assert list(r["vars"]) == VARS
    for k in VARS[:6]:
        assert r["vars"][k][0] == ("blochsteps",)
    assert r["vars"]["lambda"][0] == ("blochsteps", "eigenvalues")
    for k in VARS[9:]:
        assert r["vars"][k][0] == ("blochsteps", "eigenvalues", "graphx", "graphy")
        assert r["vars"][k][1].shape == (5, 3, 8, 6)
        assert (r["vars"][k][1] == FILL).all()
    assert np.array_equal(r["vars"]["rectangle"][1], np.array(rects).ravel())
    assert np.array_equal(r["vars"]["circle"][1], np.array(circles).ravel())
    assert (r["vars"]["lambda"][1] == FILL).all() and (r["vars"]["xbloch"][1] == FILL).all()


def test_empty_rectangle_list_is_the_unlimited_dimension(tmp_path):
    """nc_def_dim with length 0 defines the unlimited dimension: with no
    rectangles in the pattern, `rectangle` is a record variable with no
    records."""
    pat = pattern([], [(8.9, 0.5, 0.5, 0.2)])
    path = tmp_path / "b.nc"
    M.lib.del_ncfile(M.lib.new_ncfile(None, 2, 4, 4, 4, pat, str(path).encode()))
    M.lib.del_epspattern(pat)
    r = read(path)
    assert r["dims"]["rectangles"] is None and r["dims"]["circles"] == 4
    assert r["vars"]["rectangle"][1].shape == (0,)
    assert np.array_equal(r["vars"]["circle"][1], [8.9, 0.5, 0.5, 0.2])


def test_no_pattern_aborts_like_the_reference(tmp_path):
    """Without a pattern the reference defines two dimensions of length 0
    (two unlimited dimensions), which libnetcdf refuses and its assert turns
    into an abort; the same here."""
    code = ("import maxwell_amd as M; M.lib.new_ncfile(None, 1, 1, 1, 1, None, %r)" %
            str(tmp_path / "c.nc").encode())
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))), timeout=120)
    assert p.returncode != 0 and "unlimited" in p.stderr


def test_oracle_sampling_special_cases():
    """Constant fields sample to themselves; a field linear in the column
    index (row index) samples to its value plus off times the slope, with off
    the distance to the left (lower) edge in units of length -- the
    reference's weights; at the last column the Bloch factor continues it."""
    nx, ny, gx, gy = 8, 6, 5, 4
    h = 1.0 / nx, 1.0 / ny
    v = np.concatenate([np.full(nx * ny, 2 - 1j), np.full(nx * ny, 0.5 + 3j)])
    vx, vy = O.graph_sample(nx, ny, h[0], h[1], 1.0, 1.0, v, gx, gy)
    assert np.allclose(vx, 2 - 1j, atol=1e-15) and np.allclose(vy, 0.5 + 3j, atol=1e-15)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny))
    v = np.concatenate([(ii + 10 * jj).ravel(), (3 * jj + ii).ravel()]).astype(complex)
    xf = np.exp(0.3j)
    vx, vy = O.graph_sample(nx, ny, h[0], h[1], xf, 1.0, v, gx, gy)
    for ig in range(gx):
        i = int((ig + 0.5) * nx / gx)
        off = (ig + 0.5) / gx - i / nx
        for jg in range(gy):
            j = int((jg + 0.5) * ny / gy)
            nxt = (i + 1 + 10 * j) if i + 1 < nx else xf * (10 * j)
            assert abs(vx[ig, jg] - ((1 - off) * (i + 10 * j) + off * nxt)) < 1e-12


@pytest.mark.gpu
def test_write_ncfile_from_device_vectors(tmp_path):
    n, eigs, steps, gx, gy = 16, 3, 4, 10, 7
    gr = M.lib.new_grid2d(n, n, 1.0 / n, 1.0 / n)
    pat = pattern([], [(8.9, 0.5, 0.5, 0.2)])
    M.lib.setpattern_grid2d(pat, gr)
    vecs = [M.lib.new_edge2d(gr) for _ in range(eigs)]
    path = tmp_path / "w.nc"
    print("new_ncfile", path, flush=True)
    nf = M.lib.new_ncfile(vecs[0], eigs, steps, gx, gy, pat, str(path).encode())
    rng = np.random.default_rng(4)
    written = {}
    for step, (kx, ky) in ((0, (0.0, 0.0)), (2, (np.pi, 0.4))):
        g = gr.contents
        g.xbloch, g.ybloch = kx, ky
        g.xfactor = M.Cplx(np.cos(kx), np.sin(kx))
        g.yfactor = M.Cplx(np.cos(ky), np.sin(ky))
        dense = [rng.standard_normal(2 * n * n) + 1j * rng.standard_normal(2 * n * n) for _ in range(eigs)]
        for v, d in zip(vecs, dense):
            M.lib.denseto_edge2d(np.ascontiguousarray(d).ctypes.data_as(M._D), v)
        lam = np.sort(rng.uniform(1.0, 50.0, eigs + 2))  # the drivers pass longer arrays
        arr = (C.c_void_p * eigs)(*vecs)
        M.lib.write_ncfile(nf, step, lam.ctypes.data_as(M._D), arr)
        written[step] = (kx, ky, lam, dense)
    M.lib.del_ncfile(nf)
    r = read(path)
    V = r["vars"]
    for step, (kx, ky, lam, dense) in written.items():
        assert V["xbloch"][1][step] == kx and V["ybloch"][1][step] == ky
        assert V["xfactorr"][1][step] == np.cos(kx) and V["xfactori"][1][step] == np.sin(kx)
        assert V["yfactorr"][1][step] == np.cos(ky) and V["yfactori"][1][step] == np.sin(ky)
        assert np.array_equal(V["lambda"][1][step], lam[:eigs] - 1.0)
        for k in range(eigs):
            vx, vy = O.graph_sample(n, n, 1.0 / n, 1.0 / n, np.exp(1j * kx), np.exp(1j * ky), dense[k], gx, gy)
            for name, ref in (("vectorxr", vx.real), ("vectorxi", vx.imag), ("vectoryr", vy.real),
                              ("vectoryi", vy.imag)):
                assert np.abs(V[name][1][step, k] - ref).max() <= 1e-14 * np.abs(ref).max(), (step, k, name)
    for name in ("lambda", "vectorxr", "xbloch"):
        assert (V[name][1][1] == FILL).all() and (V[name][1][3] == FILL).all()
    # the picture routine samples the same field
    pgm = tmp_path / "f.png"
    M.lib.cairodraw_edge2d(vecs[0], gx, gy, pat, str(pgm).encode())
    data = open(pgm, "rb").read()
    assert data.startswith(b"P5\n%d %d\n255\n" % (gx, gy)) and len(data) == len(b"P5\n%d %d\n255\n" % (gx, gy)) + gx * gy
    for v in vecs:
        M.lib.del_edge2d(v)
    M.lib.del_epspattern(pat)
    M.lib.del_grid2d(gr)
