/*
 * quadrature.c -- interpolation, L2 functionals and L2 norms of given
 * functions on edge and node vectors (edge2d.c:643-1206, node2d.c:93-385):
 * the set-up of source problems and the error measurement of the
 * reference's test programs.  The user function is a host callback, so these
 * run on the host: evaluate at the quadrature points, combine, and move the
 * coefficients to or from the device vector.  They are collectives of the
 * calling OpenMP team like the vector functions (one thread works, all
 * return the same result).
 *
 * Quadrature rules (the reference's, including its choices at the periodic
 * seam, which use the Bloch factors as stated at each rule):
 *  - edges: the field at edge midpoints.  Interpolation takes
 *    h * (tangential component) at the midpoint.  The functional of an
 *    x-edge weighs its midpoint 1/2 and the midpoints of the four y-edges
 *    that share its end nodes 1/8 each (times hy; y-edges alike).  The norm
 *    is the four-midpoint rule per cell, weight hx hy / 4 per point, of
 *    f - u_h, u_h the edge element field of the cell.
 *  - nodes: interpolation at the nodes; functional and norm with the 2 x 2
 *    Gauss rule per cell (points at (1 -+ 1/sqrt 3)/2), bilinear basis.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>

#include "edge2d.h"
#include "node2d.h"

static void *xmalloc(size_t n)
{
  void *p = malloc(n);
  if (!p) abort();
  return p;
}

/* ----------------------------------------------------------------- edges */

void interpolate_edge2d(edge2d *x, void (*func)(const real *x, field *fx, void *data), void *data)
{
#pragma omp single
  {
    const int nx = x->gr->nx, ny = x->gr->ny;
    const real hx = x->gr->hx, hy = x->gr->hy;
    field *v = (field *)xmalloc(sizeof(field) * 2 * (size_t)nx * ny), *vy = v + (size_t)nx * ny;
    real p[2];
    field f[2];
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        p[0] = (i + 0.5) * hx; /* x-edge midpoint */
        p[1] = j * hy;
        func(p, f, data);
        v[i + (size_t)j * nx] = hx * f[0];
        p[0] = i * hx; /* y-edge midpoint */
        p[1] = (j + 0.5) * hy;
        func(p, f, data);
        vy[i + (size_t)j * nx] = hy * f[1];
      }
    denseto_edge2d(v, x);
    free(v);
  }
}

/* xbuf and ybuf receive the x and y components of the function at the
   midpoints (x-edge midpoints in their x part, y-edge midpoints in their y
   part), as in the reference. */
void l2functional_edge2d(edge2d *x, edge2d *xbuf, edge2d *ybuf,
                         void (*func)(const real *x, field *fx, void *data), void *data)
{
#pragma omp single
  {
    const int nx = x->gr->nx, ny = x->gr->ny;
    const size_t N = (size_t)nx * ny;
    const real hx = x->gr->hx, hy = x->gr->hy;
    const field xf = x->gr->xfactor, yf = x->gr->yfactor, ixf = conj(xf), iyf = conj(yf);
    /* f at x-edge midpoints (fxx: x comp, fxy: y comp) and y-edge midpoints */
    field *fb = (field *)xmalloc(sizeof(field) * 4 * N);
    field *fxx = fb, *fyx = fb + N, *fxy = fb + 2 * N, *fyy = fb + 3 * N;
    real p[2];
    field f[2];
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const size_t q = i + (size_t)j * nx;
        p[0] = (i + 0.5) * hx;
        p[1] = j * hy;
        func(p, f, data);
        fxx[q] = f[0];
        fxy[q] = f[1];
        p[0] = i * hx;
        p[1] = (j + 0.5) * hy;
        func(p, f, data);
        fyx[q] = f[0];
        fyy[q] = f[1];
      }
    field *v = (field *)xmalloc(sizeof(field) * 2 * N), *vy = v + N;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        /* x-edge (i,j): y-edges (i,j), (i+1,j), (i,j-1), (i+1,j-1) */
        const int i1 = i + 1 < nx ? i + 1 : 0, jm = j > 0 ? j - 1 : ny - 1;
        const field ci = i + 1 < nx ? 1.0 : xf, cj = j > 0 ? 1.0 : iyf;
        v[i + (size_t)j * nx] =
            hy * (0.5 * fxx[i + (size_t)j * nx] + 0.125 * fyx[i + (size_t)j * nx] +
                  0.125 * ci * fyx[i1 + (size_t)j * nx] + 0.125 * cj * fyx[i + (size_t)jm * nx] +
                  0.125 * ci * cj * fyx[i1 + (size_t)jm * nx]);
        /* y-edge (i,j): x-edges (i,j), (i-1,j), (i,j+1), (i-1,j+1) */
        const int im = i > 0 ? i - 1 : nx - 1, j1 = j + 1 < ny ? j + 1 : 0;
        const field di = i > 0 ? 1.0 : ixf, dj = j + 1 < ny ? 1.0 : yf;
        vy[i + (size_t)j * nx] =
            hx * (0.5 * fyy[i + (size_t)j * nx] + 0.125 * fxy[i + (size_t)j * nx] +
                  0.125 * di * fxy[im + (size_t)j * nx] + 0.125 * dj * fxy[i + (size_t)j1 * nx] +
                  0.125 * di * dj * fxy[im + (size_t)j1 * nx]);
      }
    denseto_edge2d(v, x);
    /* buffers: [x part | y part] = (fxx | fyx) and (fxy | fyy) */
    for (size_t q = 0; q < N; q++) v[q] = fxx[q], v[N + q] = fyx[q];
    denseto_edge2d(v, xbuf);
    for (size_t q = 0; q < N; q++) v[q] = fxy[q], v[N + q] = fyy[q];
    denseto_edge2d(v, ybuf);
    free(v);
    free(fb);
  }
}

static double abs2(field z) { return creal(z) * creal(z) + cimag(z) * cimag(z); }

real l2norm_edge2d(const edge2d *x, void (*func)(const real *x, field *fx, void *data), void *data)
{
  real result = 0.0;
#pragma omp single copyprivate(result)
  {
    const int nx = x->gr->nx, ny = x->gr->ny;
    const size_t N = (size_t)nx * ny;
    const real hx = x->gr->hx, hy = x->gr->hy;
    const field xf = x->gr->xfactor, yf = x->gr->yfactor, ixf = conj(xf), iyf = conj(yf);
    field *v = (field *)xmalloc(sizeof(field) * 2 * N);
    densefrom_edge2d(x, v);
#define XE(i, j) v[(i) + (size_t)(j) * nx]
#define YE(i, j) v[N + (i) + (size_t)(j) * nx]
    double sum = 0.0;
    real p[2];
    field f[2];
    /* x-edge midpoints on the lines j = 0..ny (line ny: the top of the last
       row of cells); x component against the edge's own coefficient, y
       component against the mean of the y-edges of the cell above and of
       the cell below */
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i < nx; i++) {
        p[0] = (i + 0.5) * hx;
        p[1] = j * hy;
        if (func) func(p, f, data);
        else f[0] = f[1] = 0.0;
        const field ex = f[0] - (j < ny ? XE(i, j) : iyf * XE(i, 0)) / hx;
        const double wx = (j == 0 || j == ny) ? 0.25 : 0.5;
        sum += wx * abs2(ex) * hx * hy;
        if (j < ny) { /* cell (i, j) above */
          const field yr = i + 1 < nx ? YE(i + 1, j) : xf * YE(0, j);
          sum += 0.25 * abs2(f[1] - 0.5 * (YE(i, j) + yr) / hy) * hx * hy;
        }
        if (j > 0) { /* cell (i, j-1) below */
          const field yr = i + 1 < nx ? YE(i + 1, j - 1) : xf * YE(0, j - 1);
          sum += 0.25 * abs2(f[1] - 0.5 * (YE(i, j - 1) + yr) / hy) * hx * hy;
        }
      }
    /* y-edge midpoints on the columns i = 0..nx (column nx: the right side of
       the last column of cells); y component against the edge's own
       coefficient, x component against the mean of the x-edges of the cell
       to the right and of the cell to the left */
    for (int j = 0; j < ny; j++)
      for (int i = 0; i <= nx; i++) {
        p[0] = i * hx;
        p[1] = (j + 0.5) * hy;
        if (func) func(p, f, data);
        else f[0] = f[1] = 0.0;
        field own;
        if (i < nx) own = YE(i, j);
        else own = (j < ny - 1 ? xf : ixf) * YE(0, j); /* the reference's factors at the seam */
        const double wy = (i == 0 || i == nx) ? 0.25 : 0.5;
        sum += wy * abs2(f[1] - own / hy) * hx * hy;
        if (i < nx) { /* cell (i, j) to the right */
          const field xt = j + 1 < ny ? XE(i, j + 1) : yf * XE(i, 0);
          sum += 0.25 * abs2(f[0] - 0.5 * (XE(i, j) + xt) / hx) * hx * hy;
        }
        if (i > 0) { /* cell (i-1, j) to the left */
          const field xt = j + 1 < ny ? XE(i - 1, j + 1) : yf * XE(i - 1, 0);
          sum += 0.25 * abs2(f[0] - 0.5 * (XE(i - 1, j) + xt) / hx) * hx * hy;
        }
      }
#undef XE
#undef YE
    free(v);
    result = sqrt(sum);
  }
  return result;
}

/* ----------------------------------------------------------------- nodes */

void interpolate_node2d(node2d *x, field (*func)(const real *x, void *data), void *data)
{
#pragma omp single
  {
    const int nx = x->gr->nx, ny = x->gr->ny;
    field *v = (field *)xmalloc(sizeof(field) * (size_t)nx * ny);
    real p[2];
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        p[0] = i * x->gr->hx;
        p[1] = j * x->gr->hy;
        v[i + (size_t)j * nx] = func(p, data);
      }
    denseto_node2d(v, x);
    free(v);
  }
}

/* 2 x 2 Gauss points of the unit cell in the reference's order, and the
   bilinear basis of corner (a, b) at a point (s, t) */
static void gauss4(real s[4], real t[4])
{
  const real q0 = 0.5 * (1.0 - sqrt(1.0 / 3.0)), q1 = 0.5 * (1.0 + sqrt(1.0 / 3.0));
  s[0] = q0, t[0] = q0;
  s[1] = q1, t[1] = q0;
  s[2] = q0, t[2] = q1;
  s[3] = q1, t[3] = q1;
}

static real bilin(int a, int b, real s, real t) { return (a ? s : 1.0 - s) * (b ? t : 1.0 - t); }

/* buf receives the function at the last Gauss point of every cell, as in
   the reference. */
void l2functional_node2d(node2d *x, node2d *buf, field (*func)(const real *x, void *data), void *data)
{
#pragma omp single
  {
    const int nx = x->gr->nx, ny = x->gr->ny;
    const size_t N = (size_t)nx * ny;
    const real hx = x->gr->hx, hy = x->gr->hy;
    const field ixf = conj(x->gr->xfactor), iyf = conj(x->gr->yfactor);
    field *v = (field *)calloc(N, sizeof(field)), *last = (field *)xmalloc(sizeof(field) * N);
    if (!v) abort();
    real s[4], t[4], p[2];
    gauss4(s, t);
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int q = 0; q < 4; q++) {
          p[0] = (i + s[q]) * hx;
          p[1] = (j + t[q]) * hy;
          const field fq = func(p, data);
          if (q == 3) last[i + (size_t)j * nx] = fq;
          for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
              /* corner (i+a, j+b); across the seam the basis function is the
                 Bloch image of the one at the wrapped node */
              int ni = i + a, nj = j + b;
              field c = 1.0;
              if (ni == nx) ni = 0, c *= ixf;
              if (nj == ny) nj = 0, c *= iyf;
              v[ni + (size_t)nj * nx] += c * bilin(a, b, s[q], t[q]) * fq;
            }
        }
    for (size_t k = 0; k < N; k++) v[k] *= 0.25 * hx * hy;
    denseto_node2d(v, x);
    denseto_node2d(last, buf);
    free(v);
    free(last);
  }
}

real l2norm_node2d(const node2d *x, field (*func)(const real *x, void *data), void *data)
{
  real result = 0.0;
#pragma omp single copyprivate(result)
  {
    const int nx = x->gr->nx, ny = x->gr->ny;
    const real hx = x->gr->hx, hy = x->gr->hy;
    const field xf = x->gr->xfactor, yf = x->gr->yfactor;
    field *v = (field *)xmalloc(sizeof(field) * (size_t)nx * ny);
    densefrom_node2d(x, v);
    real s[4], t[4], p[2];
    gauss4(s, t);
    double sum = 0.0;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        field c[2][2];
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            int ni = i + a, nj = j + b;
            field f = 1.0;
            if (ni == nx) ni = 0, f *= xf;
            if (nj == ny) nj = 0, f *= yf;
            c[a][b] = f * v[ni + (size_t)nj * nx];
          }
        for (int q = 0; q < 4; q++) {
          field u = 0.0;
          for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) u += bilin(a, b, s[q], t[q]) * c[a][b];
          if (func) {
            p[0] = (i + s[q]) * hx;
            p[1] = (j + t[q]) * hy;
            u -= func(p, data);
          }
          sum += abs2(u);
        }
      }
    free(v);
    result = sqrt(sum * 0.25 * hx * hy);
  }
  return result;
}
