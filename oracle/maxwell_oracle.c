/*
 * maxwell_oracle.c -- CPU restatement (complex FP64, C11 + OpenMP) of the
 * reference's Bloch eigensolver hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by the tests (to check results), by the
 * smoke check and by the benchmark's CPU baseline leg.  The product never
 * links it.
 *
 * Every routine restates the mathematics of the reference routine it cites; no
 * reference code is reused.  Grid values outside [0,nx) x [0,ny) are reached
 * through the Bloch relation u(i+nx, j) = xf * u(i, j), u(i, j+ny) = yf * u(i, j),
 * which is how the reference's boundary variants apply the factors (e.g.
 * edge2d.c:1362-1390, 2248-2273).
 */
#include "maxwell_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

static inline cplx gxf(const og_grid *g) { return g->xf[0] + I * g->xf[1]; }
static inline cplx gyf(const og_grid *g) { return g->yf[0] + I * g->yf[1]; }

/* Wrap (i,j) into the grid and return the Bloch phase of the wrap. */
static inline cplx wrap(const og_grid *g, int *i, int *j)
{
  cplx ph = 1.0;
  if (*i < 0) { *i += g->nx; ph *= conj(gxf(g)); }
  else if (*i >= g->nx) { *i -= g->nx; ph *= gxf(g); }
  if (*j < 0) { *j += g->ny; ph *= conj(gyf(g)); }
  else if (*j >= g->ny) { *j -= g->ny; ph *= gyf(g); }
  return ph;
}

/* Unwrapped x-edge, y-edge and node values. */
static inline cplx XE(const og_grid *g, const cplx *v, int i, int j)
{
  cplx ph = wrap(g, &i, &j);
  return ph * v[i + j * g->nx];
}
static inline cplx YE(const og_grid *g, const cplx *v, int i, int j)
{
  cplx ph = wrap(g, &i, &j);
  return ph * v[(size_t)g->nx * g->ny + i + j * g->nx];
}
static inline cplx ND(const og_grid *g, const cplx *v, int i, int j)
{
  cplx ph = wrap(g, &i, &j);
  return ph * v[i + j * g->nx];
}
static inline double ETA(const og_grid *g, int i, int j)
{
  i = (i + g->nx) % g->nx;
  j = (j + g->ny) % g->ny;
  return g->eta[i + j * g->nx];
}

/* Curl of cell (i,j) in unwrapped coordinates: bottom - top - left + right
   (edge2d.c:1505-1514 writes it as x0 - x1 - y0 + y1). */
static inline cplx curl(const og_grid *g, const cplx *v, int i, int j)
{
  return XE(g, v, i, j) - XE(g, v, i, j + 1) - YE(g, v, i, j) + YE(g, v, i + 1, j);
}

/* ======================================================================
 * grid2d
 * ====================================================================== */

/* grid2d.c:171-218: 1/eps per cell, tested at the cell's lower-left corner;
   rectangles (eps, x0, y0, w, h) first, then circles (eps, cx, cy, r). */
void og_setpattern(int nx, int ny, double hx, double hy, double eps_base,
                   int nrect, const double *rect, int ncirc, const double *circ,
                   double *eta)
{
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double x = i * hx, y = j * hy;
      double v = 1.0 / eps_base;
      for (int k = 0; k < nrect; k++) {
        const double *r = rect + 5 * k;
        if (r[1] <= x && x <= r[1] + r[3] && r[2] <= y && y <= r[2] + r[4])
          v = 1.0 / r[0];
      }
      for (int k = 0; k < ncirc; k++) {
        const double *c = circ + 4 * k;
        if ((x - c[1]) * (x - c[1]) + (y - c[2]) * (y - c[2]) <= c[3] * c[3])
          v = 1.0 / c[0];
      }
      eta[i + j * nx] = v;
    }
}

/* grid2d.c:48-68: coarse 1/eps = mean of the four children's 1/eps. */
void og_coarsen(int nx, int ny, const double *ef, double *ec)
{
  int cnx = nx / 2, cny = ny / 2;
  for (int j = 0; j < cny; j++)
    for (int i = 0; i < cnx; i++)
      ec[i + j * cnx] = 0.25 * (ef[2 * i + 2 * j * nx] + ef[2 * i + 1 + 2 * j * nx] +
                                ef[2 * i + (2 * j + 1) * nx] +
                                ef[2 * i + 1 + (2 * j + 1) * nx]);
}

/* ======================================================================
 * edge2d operators
 * ====================================================================== */

/* y += (alpha K + beta M) x  (edge2d.c:1335-1650).
 * K = sum_cells eta_c c_c c_c^H / (hx hy) with c_c the cell curl; M is the
 * Nedelec mass matrix, mx*(4,1,1) along columns of x-edges and my*(4,1,1)
 * along rows of y-edges (mx = hy/(6 hx), my = hx/(6 hy)).
 * quirk != 0 reproduces the reference's eta index for the cell below an
 * x-edge in the top row and in the last column (edge2d.c:1531, 1568, 1599,
 * 1629 use eps of the cell above); it changes nothing when eta is constant
 * across those rows. */
void og_addeval(const og_grid *g, double ar, double ai, double br, double bi,
                const double *xd, double *yd, int quirk)
{
  const int nx = g->nx, ny = g->ny;
  const size_t N = (size_t)nx * ny;
  const cplx *x = (const cplx *)xd;
  cplx *y = (cplx *)yd;
  const cplx alpha = ar + I * ai, beta = br + I * bi;
  const cplx axy = alpha / g->hx / g->hy;
  const cplx mx = beta * g->hy / g->hx / 6.0;
  const cplx my = beta * g->hx / g->hy / 6.0;
  cplx *c = (cplx *)malloc(sizeof(cplx) * N);

#pragma omp parallel for
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      c[i + j * nx] = curl(g, x, i, j);

#pragma omp parallel for
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      /* x-edge (i,j): cell above (i,j) and cell below (i,j-1) */
      int ib = i, jb = j - 1;
      cplx phb = wrap(g, &ib, &jb);
      double eta_below = (quirk && j >= 1 && (j == ny - 1 || i == nx - 1))
                             ? g->eta[i + j * nx] : g->eta[ib + jb * nx];
      cplx kx = g->eta[i + j * nx] * c[i + j * nx] - eta_below * phb * c[ib + jb * nx];
      cplx mxv = 4.0 * x[i + j * nx] + XE(g, x, i, j - 1) + XE(g, x, i, j + 1);
      y[i + j * nx] += axy * kx + mx * mxv;

      /* y-edge (i,j): cell right (i,j) and cell left (i-1,j) */
      int il = i - 1, jl = j;
      cplx phl = wrap(g, &il, &jl);
      cplx ky = g->eta[il + jl * nx] * phl * c[il + jl * nx] - g->eta[i + j * nx] * c[i + j * nx];
      cplx myv = 4.0 * x[N + i + j * nx] + YE(g, x, i - 1, j) + YE(g, x, i + 1, j);
      y[N + i + j * nx] += axy * ky + my * myv;
    }
  free(c);
}

/* energyprod_edge2d (edge2d.c:1652-1959): the reference accumulates
   conj(y) . (alpha K + beta M) x and returns the conjugate of the sum. */
void og_energyprod(const og_grid *g, double ar, double ai, double br, double bi,
                   const double *xd, const double *yd, int quirk, double *out)
{
  const size_t n = 2 * (size_t)g->nx * g->ny;
  cplx *ax = (cplx *)calloc(n, sizeof(cplx));
  og_addeval(g, ar, ai, br, bi, xd, (double *)ax, quirk);
  const cplx *y = (const cplx *)yd;
  double sr = 0.0, si = 0.0;
#pragma omp parallel for reduction(+ : sr, si)
  for (size_t k = 0; k < n; k++) {
    cplx t = conj(y[k]) * ax[k];
    sr += creal(t);
    si += cimag(t);
  }
  free(ax);
  out[0] = sr;
  out[1] = -si;
}

/* massprod_edge2d (edge2d.c:1961-2027): x^H M y. */
void og_massprod(const og_grid *g, const double *xd, const double *yd, double *out)
{
  const int nx = g->nx, ny = g->ny;
  const size_t N = (size_t)nx * ny;
  const cplx *x = (const cplx *)xd, *y = (const cplx *)yd;
  const double mx = g->hy / g->hx / 6.0, my = g->hx / g->hy / 6.0;
  double sr = 0.0, si = 0.0;
#pragma omp parallel for reduction(+ : sr, si)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      cplx mxv = mx * (4.0 * x[i + j * nx] + XE(g, x, i, j - 1) + XE(g, x, i, j + 1));
      cplx myv = my * (4.0 * x[N + i + j * nx] + YE(g, x, i - 1, j) + YE(g, x, i + 1, j));
      cplx t = conj(y[i + j * nx]) * mxv + conj(y[N + i + j * nx]) * myv;
      sr += creal(t);
      si += cimag(t);
    }
  out[0] = sr;
  out[1] = -si;
}

/* dotprod_edge2d (edge2d.c:1207-1238): x^H y. */
void og_dotprod(int n, const double *xd, const double *yd, double *out)
{
  const cplx *x = (const cplx *)xd, *y = (const cplx *)yd;
  double sr = 0.0, si = 0.0;
#pragma omp parallel for reduction(+ : sr, si)
  for (int k = 0; k < n; k++) {
    cplx t = conj(x[k]) * y[k];
    sr += creal(t);
    si += cimag(t);
  }
  out[0] = sr;
  out[1] = si;
}

/* norm2_edge2d (edge2d.c:858-885). */
double og_norm2(int n, const double *xd)
{
  double s = 0.0;
#pragma omp parallel for reduction(+ : s)
  for (int k = 0; k < 2 * n; k++)
    s += xd[k] * xd[k];
  return sqrt(s);
}

/* center_edge2d (edge2d.c:1240-1333): at the Gamma point, remove the two
   constant harmonic fields, i.e. subtract the mean of the x-edge values from
   the x-edges and the mean of the y-edge values from the y-edges. */
void og_center(const og_grid *g, double *xd)
{
  const size_t N = (size_t)g->nx * g->ny;
  cplx *x = (cplx *)xd;
  cplx sx = 0.0, sy = 0.0;
  for (size_t k = 0; k < N; k++) { sx += x[k]; sy += x[N + k]; }
  sx /= (double)N;
  sy /= (double)N;
  for (size_t k = 0; k < N; k++) { x[k] -= sx; x[N + k] -= sy; }
}

/* ======================================================================
 * Arnold-Falk-Winther vertex-patch smoother
 * ====================================================================== */

/* Exact solve of the 4x4 patch system on the star of node (i,j)
 * (edge2d.c:2029-2654).  Unknowns u0 = x(i-1,j), u1 = x(i,j), u2 = y(i,j-1),
 * u3 = y(i,j); cells LL=(i-1,j-1), RL=(i,j-1), LU=(i-1,j), RU=(i,j).  The
 * residual b - (alpha K + beta M) x is exact; the patch matrix is
 * alpha K_patch + 2 beta M_patch: the reference adds the diagonal mass block
 * twice (edge2d.c:2058-2062 against the mass rows of its residual,
 * edge2d.c:2085-2096), and this restatement keeps that. */
void og_patch_solve(const og_grid *g, double alpha, double beta,
                    const double *bd, double *xd, int i, int j)
{
  const cplx *b = (const cplx *)bd;
  cplx *x = (cplx *)xd;
  const double axy = alpha / g->hx / g->hy;
  const double mx = beta * g->hy / g->hx / 6.0;
  const double my = beta * g->hx / g->hy / 6.0;

  double eLL = axy * ETA(g, i - 1, j - 1), eRL = axy * ETA(g, i, j - 1);
  double eLU = axy * ETA(g, i - 1, j), eRU = axy * ETA(g, i, j);

  cplx qLL = eLL * curl(g, x, i - 1, j - 1);
  cplx qRL = eRL * curl(g, x, i, j - 1);
  cplx qLU = eLU * curl(g, x, i - 1, j);
  cplx qRU = eRU * curl(g, x, i, j);

  cplx r[4];
  r[0] = XE(g, b, i - 1, j) - (qLU - qLL)
         - mx * (4.0 * XE(g, x, i - 1, j) + XE(g, x, i - 1, j - 1) + XE(g, x, i - 1, j + 1));
  r[1] = XE(g, b, i, j) - (qRU - qRL)
         - mx * (4.0 * XE(g, x, i, j) + XE(g, x, i, j - 1) + XE(g, x, i, j + 1));
  r[2] = YE(g, b, i, j - 1) - (qLL - qRL)
         - my * (4.0 * YE(g, x, i, j - 1) + YE(g, x, i - 1, j - 1) + YE(g, x, i + 1, j - 1));
  r[3] = YE(g, b, i, j) - (qLU - qRU)
         - my * (4.0 * YE(g, x, i, j) + YE(g, x, i - 1, j) + YE(g, x, i + 1, j));

  /* Patch matrix (real symmetric positive definite). */
  double A[4][4] = {
      {eLL + eLU + 8.0 * mx, 0.0, -eLL, eLU},
      {0.0, eRL + eRU + 8.0 * mx, eRL, -eRU},
      {-eLL, eRL, eLL + eRL + 8.0 * my, 0.0},
      {eLU, -eRU, 0.0, eLU + eRU + 8.0 * my}};

  /* Cholesky A = L L^T and two triangular solves. */
  double L[4][4] = {{0}};
  for (int c = 0; c < 4; c++) {
    double s = A[c][c];
    for (int k = 0; k < c; k++) s -= L[c][k] * L[c][k];
    L[c][c] = sqrt(s);
    for (int rr = c + 1; rr < 4; rr++) {
      double t = A[rr][c];
      for (int k = 0; k < c; k++) t -= L[rr][k] * L[c][k];
      L[rr][c] = t / L[c][c];
    }
  }
  cplx z[4];
  for (int rr = 0; rr < 4; rr++) {
    cplx t = r[rr];
    for (int k = 0; k < rr; k++) t -= L[rr][k] * z[k];
    z[rr] = t / L[rr][rr];
  }
  for (int rr = 3; rr >= 0; rr--) {
    cplx t = z[rr];
    for (int k = rr + 1; k < 4; k++) t -= L[k][rr] * z[k];
    z[rr] = t / L[rr][rr];
  }

  /* Scatter: stored value += conj(phase) * local update. */
  const size_t N = (size_t)g->nx * g->ny;
  int ii, jj;
  cplx ph;
  ii = i - 1; jj = j; ph = wrap(g, &ii, &jj); x[ii + jj * g->nx] += conj(ph) * z[0];
  ii = i; jj = j; ph = wrap(g, &ii, &jj); x[ii + jj * g->nx] += conj(ph) * z[1];
  ii = i; jj = j - 1; ph = wrap(g, &ii, &jj); x[N + ii + jj * g->nx] += conj(ph) * z[2];
  ii = i; jj = j; ph = wrap(g, &ii, &jj); x[N + ii + jj * g->nx] += conj(ph) * z[3];
}

/* One multiplicative sweep over all vertex patches.
 * colour_order 0: the reference's active line order (edge2d.c:2656-2861):
 *   forward = row 0, even rows 2..ny-2, odd rows 1..ny-3, row ny-1, each row
 *   with i ascending; backward = the exact reverse.  Rows of equal parity are
 *   independent, so they run in parallel as in the reference.
 * colour_order 1: four colours (j even, i even), (j even, i odd),
 *   (j odd, i even), (j odd, i odd) -- the inactive variant of
 *   edge2d.c:3043-3480 and the order used on the GPU; backward reverses the
 *   colours.  Patches of one colour are independent. */
void og_gs_edge(const og_grid *g, double alpha, double beta,
                const double *b, double *x, int backward, int colour_order)
{
  const int nx = g->nx, ny = g->ny;
  if (colour_order == 0) {
    if (!backward) {
      for (int i = 0; i < nx; i++) og_patch_solve(g, alpha, beta, b, x, i, 0);
#pragma omp parallel for
      for (int j = 2; j < ny - 1; j += 2)
        for (int i = 0; i < nx; i++) og_patch_solve(g, alpha, beta, b, x, i, j);
#pragma omp parallel for
      for (int j = 1; j < ny - 1; j += 2)
        for (int i = 0; i < nx; i++) og_patch_solve(g, alpha, beta, b, x, i, j);
      if (ny > 1)
        for (int i = 0; i < nx; i++) og_patch_solve(g, alpha, beta, b, x, i, ny - 1);
    } else {
      if (ny > 1)
        for (int i = nx - 1; i >= 0; i--) og_patch_solve(g, alpha, beta, b, x, i, ny - 1);
#pragma omp parallel for
      for (int j = 1; j < ny - 1; j += 2)
        for (int i = nx - 1; i >= 0; i--) og_patch_solve(g, alpha, beta, b, x, i, j);
#pragma omp parallel for
      for (int j = 2; j < ny - 1; j += 2)
        for (int i = nx - 1; i >= 0; i--) og_patch_solve(g, alpha, beta, b, x, i, j);
      for (int i = nx - 1; i >= 0; i--) og_patch_solve(g, alpha, beta, b, x, i, 0);
    }
  } else {
    for (int s = 0; s < 4; s++) {
      int c = backward ? 3 - s : s;
      int pi = c & 1, pj = c >> 1;
#pragma omp parallel for
      for (int j = pj; j < ny; j += 2)
        for (int i = pi; i < nx; i += 2) og_patch_solve(g, alpha, beta, b, x, i, j);
    }
  }
}

/* ======================================================================
 * edge transfers (edge2d.c:5583-5705)
 * ====================================================================== */

/* f += alpha P c.  A coarse edge is split into two collinear fine edges with
   weight 1/2 each (the DOF is a line integral) and feeds the fine edge
   between it and its parallel neighbour with weight 1/4. */
void og_prolong_edge(const og_grid *gf, double ar, double ai, const double *cd, double *fd)
{
  const int fnx = gf->nx, fny = gf->ny, cnx = fnx / 2, cny = fny / 2;
  const size_t FN = (size_t)fnx * fny;
  og_grid gc = *gf;
  gc.nx = cnx; gc.ny = cny; gc.hx = 2 * gf->hx; gc.hy = 2 * gf->hy; gc.eta = 0;
  const cplx *c = (const cplx *)cd;
  cplx *f = (cplx *)fd;
  const cplx a = ar + I * ai;
#pragma omp parallel for
  for (int j = 0; j < cny; j++)
    for (int i = 0; i < cnx; i++) {
      cplx cx0 = XE(&gc, c, i, j), cx1 = XE(&gc, c, i, j + 1);
      f[2 * i + 2 * j * fnx] += a * 0.5 * cx0;
      f[2 * i + 1 + 2 * j * fnx] += a * 0.5 * cx0;
      f[2 * i + (2 * j + 1) * fnx] += a * 0.25 * (cx0 + cx1);
      f[2 * i + 1 + (2 * j + 1) * fnx] += a * 0.25 * (cx0 + cx1);
      cplx cy0 = YE(&gc, c, i, j), cy1 = YE(&gc, c, i + 1, j);
      f[FN + 2 * i + 2 * j * fnx] += a * 0.5 * cy0;
      f[FN + 2 * i + (2 * j + 1) * fnx] += a * 0.5 * cy0;
      f[FN + 2 * i + 1 + 2 * j * fnx] += a * 0.25 * (cy0 + cy1);
      f[FN + 2 * i + 1 + (2 * j + 1) * fnx] += a * 0.25 * (cy0 + cy1);
    }
}

/* c += alpha P^H f. */
void og_restrict_edge(const og_grid *gf, double ar, double ai, const double *fd, double *cd)
{
  const int fnx = gf->nx, fny = gf->ny, cnx = fnx / 2, cny = fny / 2;
  const size_t CN = (size_t)cnx * cny;
  const cplx *f = (const cplx *)fd;
  cplx *c = (cplx *)cd;
  const cplx a = ar + I * ai;
#pragma omp parallel for
  for (int j = 0; j < cny; j++)
    for (int i = 0; i < cnx; i++) {
      cplx sx = 0.5 * (XE(gf, f, 2 * i, 2 * j) + XE(gf, f, 2 * i + 1, 2 * j))
                + 0.25 * (XE(gf, f, 2 * i, 2 * j - 1) + XE(gf, f, 2 * i + 1, 2 * j - 1)
                          + XE(gf, f, 2 * i, 2 * j + 1) + XE(gf, f, 2 * i + 1, 2 * j + 1));
      cplx sy = 0.5 * (YE(gf, f, 2 * i, 2 * j) + YE(gf, f, 2 * i, 2 * j + 1))
                + 0.25 * (YE(gf, f, 2 * i - 1, 2 * j) + YE(gf, f, 2 * i - 1, 2 * j + 1)
                          + YE(gf, f, 2 * i + 1, 2 * j) + YE(gf, f, 2 * i + 1, 2 * j + 1));
      c[i + j * cnx] += a * sx;
      c[CN + i + j * cnx] += a * sy;
    }
}

/* ======================================================================
 * dense assembly (edge2d.c:5870-6110, node2d.c:1429-1602) and small dense
 * linear algebra (the LAPACK calls of linalg.c:53-63, 98-113, 291-317)
 * ====================================================================== */

/* a (column-major, n x n with n = 2 nx ny) = alpha K + beta M, assembled
   cell by cell from the element matrices, an independent route to the same
   matrix as the stencil. */
void og_dense_edge(const og_grid *g, double ar, double ai, double br, double bi, double *ad)
{
  const int nx = g->nx, ny = g->ny;
  const int N = nx * ny, n = 2 * N;
  cplx *a = (cplx *)ad;
  const cplx alpha = ar + I * ai, beta = br + I * bi;
  const cplx axy = alpha / g->hx / g->hy;
  const cplx mx = beta * g->hy / g->hx / 6.0;
  const cplx my = beta * g->hx / g->hy / 6.0;
  memset(a, 0, sizeof(cplx) * (size_t)n * n);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      /* local DOFs: bottom x(i,j), top x(i,j+1), left y(i,j), right y(i+1,j) */
      int gi[4];
      cplx ph[4];
      int ii, jj;
      ii = i; jj = j; ph[0] = wrap(g, &ii, &jj); gi[0] = ii + jj * nx;
      ii = i; jj = j + 1; ph[1] = wrap(g, &ii, &jj); gi[1] = ii + jj * nx;
      ii = i; jj = j; ph[2] = wrap(g, &ii, &jj); gi[2] = N + ii + jj * nx;
      ii = i + 1; jj = j; ph[3] = wrap(g, &ii, &jj); gi[3] = N + ii + jj * nx;
      const double v[4] = {1.0, -1.0, -1.0, 1.0};
      cplx ke[4][4];
      for (int p = 0; p < 4; p++)
        for (int q = 0; q < 4; q++) ke[p][q] = axy * g->eta[i + j * nx] * v[p] * v[q];
      ke[0][0] += 2.0 * mx; ke[1][1] += 2.0 * mx; ke[0][1] += mx; ke[1][0] += mx;
      ke[2][2] += 2.0 * my; ke[3][3] += 2.0 * my; ke[2][3] += my; ke[3][2] += my;
      for (int p = 0; p < 4; p++)
        for (int q = 0; q < 4; q++)
          a[gi[p] + (size_t)gi[q] * n] += conj(ph[p]) * ke[p][q] * ph[q];
    }
}

/* Q1 nodal Laplacian (node2d.c:1429-1602) by element assembly; with fixnull
   at the Gamma point every entry gets 0.1/(n*n) (node2d.c:1454-1458). */
void og_dense_node(const og_grid *g, int fixnull, double *ad)
{
  const int nx = g->nx, ny = g->ny, n = nx * ny;
  cplx *a = (cplx *)ad;
  const double rx = g->hy / g->hx, ry = g->hx / g->hy;
  double fill = 0.0;
  if (fixnull && cabs(gxf(g) - 1.0) < 1e-8 && cabs(gyf(g) - 1.0) < 1e-8)
    fill = 0.1 / ((double)n * n);
  for (size_t k = 0; k < (size_t)n * n; k++) a[k] = fill;
  /* 1-D stiffness S and mass T on the reference interval */
  const double S[2][2] = {{1.0, -1.0}, {-1.0, 1.0}};
  const double T[2][2] = {{1.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 3.0}};
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int gi[4];
      cplx ph[4];
      for (int p = 0; p < 4; p++) {
        int ii = i + (p & 1), jj = j + (p >> 1);
        ph[p] = wrap(g, &ii, &jj);
        gi[p] = ii + jj * nx;
      }
      for (int p = 0; p < 4; p++)
        for (int q = 0; q < 4; q++) {
          int px = p & 1, py = p >> 1, qx = q & 1, qy = q >> 1;
          double ke = rx * S[px][qx] * T[py][qy] + ry * T[px][qx] * S[py][qy];
          a[gi[p] + (size_t)gi[q] * n] += conj(ph[p]) * ke * ph[q];
        }
    }
}

/* Lower Cholesky a = L L^H in place (column-major).  Returns 0 or the
   1-based index of the first non-positive pivot, as zpotrf does. */
int og_potrf(int n, double *ad)
{
  cplx *a = (cplx *)ad;
  for (int c = 0; c < n; c++) {
    double s = creal(a[c + (size_t)c * n]);
    for (int k = 0; k < c; k++) s -= creal(a[c + (size_t)k * n] * conj(a[c + (size_t)k * n]));
    if (!(s > 0.0)) return c + 1;
    double d = sqrt(s);
    a[c + (size_t)c * n] = d;
    for (int r = c + 1; r < n; r++) {
      cplx t = a[r + (size_t)c * n];
      for (int k = 0; k < c; k++) t -= a[r + (size_t)k * n] * conj(a[c + (size_t)k * n]);
      a[r + (size_t)c * n] = t / d;
    }
    for (int r = 0; r < c; r++) a[r + (size_t)c * n] = 0.0;
  }
  return 0;
}

/* Solve L L^H x = b in place (potrsv_matrix, linalg.c:98-113). */
void og_potrs(int n, const double *ld, double *bd)
{
  const cplx *l = (const cplx *)ld;
  cplx *b = (cplx *)bd;
  for (int r = 0; r < n; r++) {
    cplx t = b[r];
    for (int k = 0; k < r; k++) t -= l[r + (size_t)k * n] * b[k];
    b[r] = t / l[r + (size_t)r * n];
  }
  for (int r = n - 1; r >= 0; r--) {
    cplx t = b[r];
    for (int k = r + 1; k < n; k++) t -= conj(l[k + (size_t)r * n]) * b[k];
    b[r] = t / conj(l[r + (size_t)r * n]);
  }
}

/* Cyclic Jacobi for a full Hermitian matrix c (n x n, column-major);
   eigenvalues to w (ascending), eigenvectors to the columns of v. */
static void jacobi_heev(int n, cplx *c, double *w, cplx *v)
{
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) v[i + (size_t)j * n] = (i == j);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        double t = creal(c[i + (size_t)j * n] * conj(c[i + (size_t)j * n]));
        if (i == j) dia += t; else off += t;
      }
    if (off <= 1e-32 * dia || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        cplx z = c[p + (size_t)q * n];
        double r = cabs(z);
        if (r == 0.0) continue;
        double app = creal(c[p + (size_t)p * n]), aqq = creal(c[q + (size_t)q * n]);
        double tau = (aqq - app) / (2.0 * r);
        double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        cplx e = conj(z) / r; /* e^{-i phi} */
        /* U = [[cs, sn], [-sn e, cs e]] on columns p, q */
        for (int k = 0; k < n; k++) {
          cplx ckp = c[k + (size_t)p * n], ckq = c[k + (size_t)q * n];
          c[k + (size_t)p * n] = cs * ckp - sn * e * ckq;
          c[k + (size_t)q * n] = sn * ckp + cs * e * ckq;
        }
        for (int k = 0; k < n; k++) {
          cplx cpk = c[p + (size_t)k * n], cqk = c[q + (size_t)k * n];
          c[p + (size_t)k * n] = cs * cpk - sn * conj(e) * cqk;
          c[q + (size_t)k * n] = sn * cpk + cs * conj(e) * cqk;
        }
        c[p + (size_t)q * n] = 0.0;
        c[q + (size_t)p * n] = 0.0;
        c[p + (size_t)p * n] = creal(c[p + (size_t)p * n]);
        c[q + (size_t)q * n] = creal(c[q + (size_t)q * n]);
        for (int k = 0; k < n; k++) {
          cplx vkp = v[k + (size_t)p * n], vkq = v[k + (size_t)q * n];
          v[k + (size_t)p * n] = cs * vkp - sn * e * vkq;
          v[k + (size_t)q * n] = sn * vkp + cs * e * vkq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = creal(c[i + (size_t)i * n]);
  /* selection sort, ascending */
  for (int i = 0; i < n; i++) {
    int m = i;
    for (int k = i + 1; k < n; k++) if (w[k] < w[m]) m = k;
    if (m != i) {
      double t = w[i]; w[i] = w[m]; w[m] = t;
      for (int k = 0; k < n; k++) {
        cplx s = v[k + (size_t)i * n];
        v[k + (size_t)i * n] = v[k + (size_t)m * n];
        v[k + (size_t)m * n] = s;
      }
    }
  }
}

/* Generalised Hermitian-definite eigenproblem A x = w B x (zhegv itype 1 with
   uplo "L", linalg.c:291-317): lower triangles of a and b are read; on
   return a holds B-orthonormal eigenvectors, b its Cholesky factor. */
int og_hegv(int n, double *ad, double *bd, double *w)
{
  cplx *a = (cplx *)ad, *b = (cplx *)bd;
  if (n == 0) return 0;
  int info = og_potrf(n, bd);
  if (info) return n + info;
  cplx *c = (cplx *)malloc(sizeof(cplx) * (size_t)n * n);
  cplx *v = (cplx *)malloc(sizeof(cplx) * (size_t)n * n);
  /* full Hermitian A from its lower triangle */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      c[i + (size_t)j * n] = (i >= j) ? a[i + (size_t)j * n] : conj(a[j + (size_t)i * n]);
  for (int i = 0; i < n; i++) c[i + (size_t)i * n] = creal(c[i + (size_t)i * n]);
  /* c <- L^{-1} c L^{-H}: solve on columns, then on rows */
  for (int j = 0; j < n; j++)
    for (int r = 0; r < n; r++) {
      cplx t = c[r + (size_t)j * n];
      for (int k = 0; k < r; k++) t -= b[r + (size_t)k * n] * c[k + (size_t)j * n];
      c[r + (size_t)j * n] = t / b[r + (size_t)r * n];
    }
  for (int i = 0; i < n; i++)
    for (int r = 0; r < n; r++) {
      /* row i of c times L^{-H}: solve x L^H = row, i.e. L conj(x)^T = conj(row)^T */
      cplx t = c[i + (size_t)r * n];
      for (int k = 0; k < r; k++) t -= c[i + (size_t)k * n] * conj(b[r + (size_t)k * n]);
      c[i + (size_t)r * n] = t / conj(b[r + (size_t)r * n]);
    }
  jacobi_heev(n, c, w, v);
  /* x = L^{-H} y */
  for (int j = 0; j < n; j++)
    for (int r = n - 1; r >= 0; r--) {
      cplx t = v[r + (size_t)j * n];
      for (int k = r + 1; k < n; k++) t -= conj(b[k + (size_t)r * n]) * v[k + (size_t)j * n];
      v[r + (size_t)j * n] = t / conj(b[r + (size_t)r * n]);
    }
  memcpy(a, v, sizeof(cplx) * (size_t)n * n);
  free(v);
  free(c);
  return 0;
}

/* ======================================================================
 * node2d
 * ====================================================================== */

/* y += alpha A x, A the Q1 nine-point Laplacian (node2d.c:516-627). */
void og_addeval_node(const og_grid *g, double ar, double ai, const double *xd, double *yd)
{
  const int nx = g->nx, ny = g->ny;
  const cplx *x = (const cplx *)xd;
  cplx *y = (cplx *)yd;
  const cplx al = ar + I * ai;
  const double hx = g->hx, hy = g->hy;
  const cplx ac = 4.0 * al * (hx / hy + hy / hx) / 3.0;
  const cplx ax = al * (hx / hy - 2.0 * hy / hx) / 3.0;
  const cplx ay = al * (hy / hx - 2.0 * hx / hy) / 3.0;
  const cplx adg = -al * (hx / hy + hy / hx) / 6.0;
#pragma omp parallel for
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      y[i + j * nx] += ac * x[i + j * nx]
                       + ax * (ND(g, x, i - 1, j) + ND(g, x, i + 1, j))
                       + ay * (ND(g, x, i, j - 1) + ND(g, x, i, j + 1))
                       + adg * (ND(g, x, i - 1, j - 1) + ND(g, x, i + 1, j - 1)
                                + ND(g, x, i - 1, j + 1) + ND(g, x, i + 1, j + 1));
}

static inline void node_update(const og_grid *g, double al, const cplx *b, cplx *x, int i, int j)
{
  const double hx = g->hx, hy = g->hy;
  const double ac = 4.0 * al * (hx / hy + hy / hx) / 3.0;
  const double ax = al * (hx / hy - 2.0 * hy / hx) / 3.0;
  const double ay = al * (hy / hx - 2.0 * hx / hy) / 3.0;
  const double adg = -al * (hx / hy + hy / hx) / 6.0;
  x[i + j * g->nx] = (b[i + j * g->nx]
                      - ax * (ND(g, x, i - 1, j) + ND(g, x, i + 1, j))
                      - ay * (ND(g, x, i, j - 1) + ND(g, x, i, j + 1))
                      - adg * (ND(g, x, i - 1, j - 1) + ND(g, x, i + 1, j - 1)
                               + ND(g, x, i - 1, j + 1) + ND(g, x, i + 1, j + 1))) / ac;
}

/* Pointwise Gauss-Seidel.  colour_order 0: the reference's line order
   (node2d.c:629-877): forward = row 0, odd rows, even rows, row ny-1, i
   ascending; backward = exact reverse.  colour_order 1: four colours by
   (i mod 2, j mod 2), the GPU order. */
void og_gs_node(const og_grid *g, double alpha, const double *bd, double *xd,
                int backward, int colour_order)
{
  const int nx = g->nx, ny = g->ny;
  const cplx *b = (const cplx *)bd;
  cplx *x = (cplx *)xd;
  if (colour_order == 0) {
    if (!backward) {
      for (int i = 0; i < nx; i++) node_update(g, alpha, b, x, i, 0);
#pragma omp parallel for
      for (int j = 1; j < ny - 1; j += 2)
        for (int i = 0; i < nx; i++) node_update(g, alpha, b, x, i, j);
#pragma omp parallel for
      for (int j = 2; j < ny - 1; j += 2)
        for (int i = 0; i < nx; i++) node_update(g, alpha, b, x, i, j);
      if (ny > 1)
        for (int i = 0; i < nx; i++) node_update(g, alpha, b, x, i, ny - 1);
    } else {
      if (ny > 1)
        for (int i = nx - 1; i >= 0; i--) node_update(g, alpha, b, x, i, ny - 1);
#pragma omp parallel for
      for (int j = 2; j < ny - 1; j += 2)
        for (int i = nx - 1; i >= 0; i--) node_update(g, alpha, b, x, i, j);
#pragma omp parallel for
      for (int j = 1; j < ny - 1; j += 2)
        for (int i = nx - 1; i >= 0; i--) node_update(g, alpha, b, x, i, j);
      for (int i = nx - 1; i >= 0; i--) node_update(g, alpha, b, x, i, 0);
    }
  } else {
    for (int s = 0; s < 4; s++) {
      int c = backward ? 3 - s : s;
      int pi = c & 1, pj = c >> 1;
#pragma omp parallel for
      for (int j = pj; j < ny; j += 2)
        for (int i = pi; i < nx; i += 2) node_update(g, alpha, b, x, i, j);
    }
  }
}

/* Bilinear prolongation f += alpha P c (node2d.c:1140-1207). */
void og_prolong_node(const og_grid *gf, double ar, double ai, const double *cd, double *fd)
{
  const int fnx = gf->nx, cnx = fnx / 2, cny = gf->ny / 2;
  og_grid gc = *gf;
  gc.nx = cnx; gc.ny = cny; gc.eta = 0;
  const cplx *c = (const cplx *)cd;
  cplx *f = (cplx *)fd;
  const cplx a = ar + I * ai;
#pragma omp parallel for
  for (int j = 0; j < cny; j++)
    for (int i = 0; i < cnx; i++) {
      cplx c00 = c[i + j * cnx], c10 = ND(&gc, c, i + 1, j);
      cplx c01 = ND(&gc, c, i, j + 1), c11 = ND(&gc, c, i + 1, j + 1);
      f[2 * i + 2 * j * fnx] += a * c00;
      f[2 * i + 1 + 2 * j * fnx] += a * 0.5 * (c00 + c10);
      f[2 * i + (2 * j + 1) * fnx] += a * 0.5 * (c00 + c01);
      f[2 * i + 1 + (2 * j + 1) * fnx] += a * 0.25 * (c00 + c10 + c01 + c11);
    }
}

/* c += alpha P^H f (node2d.c:1209-1273). */
void og_restrict_node(const og_grid *gf, double ar, double ai, const double *fd, double *cd)
{
  const int cnx = gf->nx / 2, cny = gf->ny / 2;
  const cplx *f = (const cplx *)fd;
  cplx *c = (cplx *)cd;
  const cplx a = ar + I * ai;
  const double w[3] = {0.5, 1.0, 0.5};
#pragma omp parallel for
  for (int j = 0; j < cny; j++)
    for (int i = 0; i < cnx; i++) {
      cplx s = 0.0;
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
          s += w[di + 1] * w[dj + 1] * ND(gf, f, 2 * i + di, 2 * j + dj);
      c[i + j * cnx] += a * s;
    }
}

/* e += alpha G v: edge value = difference of its end nodes (node2d.c:1275-1310). */
void og_gradient(const og_grid *g, double ar, double ai, const double *vd, double *ed)
{
  const int nx = g->nx, ny = g->ny;
  const size_t N = (size_t)nx * ny;
  const cplx *v = (const cplx *)vd;
  cplx *e = (cplx *)ed;
  const cplx a = ar + I * ai;
#pragma omp parallel for
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      e[i + j * nx] += a * (ND(g, v, i + 1, j) - v[i + j * nx]);
      e[N + i + j * nx] += a * (ND(g, v, i, j + 1) - v[i + j * nx]);
    }
}

/* v += alpha G^H e (node2d.c:1312-1350). */
void og_adjgradient(const og_grid *g, double ar, double ai, const double *ed, double *vd)
{
  const int nx = g->nx, ny = g->ny;
  const cplx *e = (const cplx *)ed;
  cplx *v = (cplx *)vd;
  const cplx a = ar + I * ai;
#pragma omp parallel for
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      v[i + j * nx] += a * (XE(g, e, i - 1, j) - e[i + j * nx]
                            + YE(g, e, i, j - 1) - e[(size_t)nx * ny + i + j * nx]);
}

/* ======================================================================
 * multigrid hierarchy, V-cycles, PINVIT (edge2d.c:5707-5754,
 * node2d.c:1352-1389, edge2d.c:6298-6328, exp_bloch2d.c:956-1053)
 * ====================================================================== */

#define OG_MAXL 16

struct og_mg {
  int n0, L, colour_order, quirk;
  og_grid g[OG_MAXL];
  double *eta[OG_MAXL];
  cplx *b[OG_MAXL], *x[OG_MAXL], *d[OG_MAXL];
  cplx *bg[OG_MAXL], *xg[OG_MAXL], *dg[OG_MAXL];
  cplx *Ae, *An; /* Cholesky factors of the coarse edge (K+M) and node matrices */
};

og_mg *og_mg_new(int n0, int L, const double *eta_fine, int colour_order, int quirk)
{
  if (L >= OG_MAXL) return NULL;
  og_mg *mg = (og_mg *)calloc(1, sizeof(og_mg));
  mg->n0 = n0; mg->L = L; mg->colour_order = colour_order; mg->quirk = quirk;
  for (int l = 0; l <= L; l++) {
    int n = n0 << l;
    size_t N = (size_t)n * n;
    mg->g[l].nx = n; mg->g[l].ny = n;
    mg->g[l].hx = 1.0 / n; mg->g[l].hy = 1.0 / n;
    mg->g[l].xf[0] = 1.0; mg->g[l].yf[0] = 1.0;
    mg->eta[l] = (double *)malloc(sizeof(double) * N);
    mg->g[l].eta = mg->eta[l];
    mg->b[l] = (cplx *)calloc(2 * N, sizeof(cplx));
    mg->x[l] = (cplx *)calloc(2 * N, sizeof(cplx));
    mg->d[l] = (cplx *)calloc(2 * N, sizeof(cplx));
    mg->bg[l] = (cplx *)calloc(N, sizeof(cplx));
    mg->xg[l] = (cplx *)calloc(N, sizeof(cplx));
    mg->dg[l] = (cplx *)calloc(N, sizeof(cplx));
  }
  memcpy(mg->eta[L], eta_fine, sizeof(double) * (size_t)(n0 << L) * (n0 << L));
  for (int l = L; l > 0; l--) og_coarsen(n0 << l, n0 << l, mg->eta[l], mg->eta[l - 1]);
  int ne = 2 * n0 * n0, nn = n0 * n0;
  mg->Ae = (cplx *)malloc(sizeof(cplx) * ne * ne);
  mg->An = (cplx *)malloc(sizeof(cplx) * nn * nn);
  og_mg_set_bloch(mg, 0.0, 0.0);
  return mg;
}

void og_mg_free(og_mg *mg)
{
  if (!mg) return;
  for (int l = 0; l <= mg->L; l++) {
    free(mg->eta[l]); free(mg->b[l]); free(mg->x[l]); free(mg->d[l]);
    free(mg->bg[l]); free(mg->xg[l]); free(mg->dg[l]);
  }
  free(mg->Ae); free(mg->An); free(mg);
}

const double *og_mg_eta(const og_mg *mg, int l) { return mg->eta[l]; }

/* Bloch factors on every level and the per-k coarse rebuild
   (exp_bloch2d.c:786-809). */
void og_mg_set_bloch(og_mg *mg, double kx, double ky)
{
  for (int l = 0; l <= mg->L; l++) {
    mg->g[l].xf[0] = cos(kx); mg->g[l].xf[1] = sin(kx);
    mg->g[l].yf[0] = cos(ky); mg->g[l].yf[1] = sin(ky);
  }
  int ne = 2 * mg->n0 * mg->n0, nn = mg->n0 * mg->n0;
  og_dense_edge(&mg->g[0], 1.0, 0.0, 1.0, 0.0, (double *)mg->Ae);
  og_potrf(ne, (double *)mg->Ae);
  og_dense_node(&mg->g[0], 1, (double *)mg->An);
  og_potrf(nn, (double *)mg->An);
}

static size_t edim(const og_mg *mg, int l) { return 2 * (size_t)mg->g[l].nx * mg->g[l].ny; }
static size_t ndim(const og_mg *mg, int l) { return (size_t)mg->g[l].nx * mg->g[l].ny; }

/* V(nu,nu) on the edge hierarchy with the coarse Cholesky solve. */
static void vcycle_edge_levels(og_mg *mg, int nu, cplx **b, cplx **x)
{
  const int L = mg->L;
  for (int l = L; l > 0; l--) {
    for (int s = 0; s < nu; s++)
      og_gs_edge(&mg->g[l], 1.0, 1.0, (double *)b[l], (double *)x[l], 0, mg->colour_order);
    memcpy(mg->d[l], b[l], sizeof(cplx) * edim(mg, l));
    og_addeval(&mg->g[l], -1.0, 0.0, -1.0, 0.0, (double *)x[l], (double *)mg->d[l], mg->quirk);
    memset(b[l - 1], 0, sizeof(cplx) * edim(mg, l - 1));
    og_restrict_edge(&mg->g[l], 1.0, 0.0, (double *)mg->d[l], (double *)b[l - 1]);
    memset(x[l - 1], 0, sizeof(cplx) * edim(mg, l - 1));
  }
  memcpy(x[0], b[0], sizeof(cplx) * edim(mg, 0));
  og_potrs((int)edim(mg, 0), (const double *)mg->Ae, (double *)x[0]);
  for (int l = 1; l <= L; l++) {
    og_prolong_edge(&mg->g[l], 1.0, 0.0, (const double *)x[l - 1], (double *)x[l]);
    for (int s = 0; s < nu; s++)
      og_gs_edge(&mg->g[l], 1.0, 1.0, (double *)b[l], (double *)x[l], 1, mg->colour_order);
  }
}

void og_mg_vcycle_edge(og_mg *mg, int nu, const double *b, double *x)
{
  cplx *bb[OG_MAXL], *xx[OG_MAXL];
  for (int l = 0; l < mg->L; l++) { bb[l] = mg->b[l]; xx[l] = mg->x[l]; }
  bb[mg->L] = (cplx *)b; xx[mg->L] = (cplx *)x;
  vcycle_edge_levels(mg, nu, bb, xx);
}

static void vcycle_node_levels(og_mg *mg, int nu, cplx **b, cplx **x)
{
  const int L = mg->L;
  for (int l = L; l > 0; l--) {
    for (int s = 0; s < nu; s++)
      og_gs_node(&mg->g[l], 1.0, (double *)b[l], (double *)x[l], 0, mg->colour_order);
    memcpy(mg->dg[l], b[l], sizeof(cplx) * ndim(mg, l));
    og_addeval_node(&mg->g[l], -1.0, 0.0, (double *)x[l], (double *)mg->dg[l]);
    memset(b[l - 1], 0, sizeof(cplx) * ndim(mg, l - 1));
    og_restrict_node(&mg->g[l], 1.0, 0.0, (double *)mg->dg[l], (double *)b[l - 1]);
    memset(x[l - 1], 0, sizeof(cplx) * ndim(mg, l - 1));
  }
  memcpy(x[0], b[0], sizeof(cplx) * ndim(mg, 0));
  og_potrs((int)ndim(mg, 0), (const double *)mg->An, (double *)x[0]);
  for (int l = 1; l <= L; l++) {
    og_prolong_node(&mg->g[l], 1.0, 0.0, (const double *)x[l - 1], (double *)x[l]);
    for (int s = 0; s < nu; s++)
      og_gs_node(&mg->g[l], 1.0, (double *)b[l], (double *)x[l], 1, mg->colour_order);
  }
}

void og_mg_vcycle_node(og_mg *mg, int nu, const double *b, double *x)
{
  cplx *bb[OG_MAXL], *xx[OG_MAXL];
  for (int l = 0; l < mg->L; l++) { bb[l] = mg->bg[l]; xx[l] = mg->xg[l]; }
  bb[mg->L] = (cplx *)b; xx[mg->L] = (cplx *)x;
  vcycle_node_levels(mg, nu, bb, xx);
}

static int is_gamma(const og_grid *g)
{
  return g->xf[0] == 1.0 && g->xf[1] == 0.0 && g->yf[0] == 1.0 && g->yf[1] == 0.0;
}

void og_lift(og_mg *mg, int gradient_steps, double *e)
{
  const int L = mg->L;
  const og_grid *g = &mg->g[L];
  memset(mg->d[L], 0, sizeof(cplx) * edim(mg, L));
  og_addeval(g, 0.0, 0.0, 1.0, 0.0, e, (double *)mg->d[L], mg->quirk);
  memset(mg->bg[L], 0, sizeof(cplx) * ndim(mg, L));
  og_adjgradient(g, 1.0, 0.0, (const double *)mg->d[L], (double *)mg->bg[L]);
  memset(mg->xg[L], 0, sizeof(cplx) * ndim(mg, L));
  for (int s = 0; s < gradient_steps; s++)
    vcycle_node_levels(mg, 2, mg->bg, mg->xg);
  og_gradient(g, -1.0, 0.0, (const double *)mg->xg[L], e);
}

void og_pinvit_vector(og_mg *mg, int prec_steps, int gradient_steps, double lam, double *e)
{
  const int L = mg->L;
  const og_grid *g = &mg->g[L];
  const size_t n = edim(mg, L);
  memset(mg->b[L], 0, sizeof(cplx) * n);
  og_addeval(g, 1.0, 0.0, 1.0 - lam, 0.0, e, (double *)mg->b[L], mg->quirk);
  memset(mg->x[L], 0, sizeof(cplx) * n);
  for (int s = 0; s < prec_steps; s++)
    vcycle_edge_levels(mg, 2, mg->b, mg->x);
  cplx *ec = (cplx *)e;
  for (size_t k = 0; k < n; k++) ec[k] -= mg->x[L][k];
  if (is_gamma(g)) og_center(g, e);
  og_lift(mg, gradient_steps, e);
}

/* Rayleigh-Ritz on the block: Ar = U^H (K+M) U, Mr = U^H M U (lower
   triangles, exp_bloch2d.c:992-996), hegv, E = U C (exp_bloch2d.c:1006-1010). */
static int ritz(og_mg *mg, int blk, cplx *E, double *w)
{
  const og_grid *g = &mg->g[mg->L];
  const size_t n = edim(mg, mg->L);
  cplx *Ar = (cplx *)calloc((size_t)blk * blk, sizeof(cplx));
  cplx *Mr = (cplx *)calloc((size_t)blk * blk, sizeof(cplx));
  for (int j = 0; j < blk; j++)
    for (int i = j; i < blk; i++) {
      double v[2];
      og_energyprod(g, 1.0, 0.0, 1.0, 0.0, (const double *)(E + i * n),
                    (const double *)(E + j * n), mg->quirk, v);
      Ar[i + j * blk] = v[0] + I * v[1];
      og_massprod(g, (const double *)(E + i * n), (const double *)(E + j * n), v);
      Mr[i + j * blk] = v[0] + I * v[1];
    }
  int info = og_hegv(blk, (double *)Ar, (double *)Mr, w);
  if (info) { free(Ar); free(Mr); return info; }
  cplx *U = (cplx *)malloc(sizeof(cplx) * n * blk);
  memcpy(U, E, sizeof(cplx) * n * blk);
#pragma omp parallel for
  for (size_t k = 0; k < n; k++)
    for (int j = 0; j < blk; j++) {
      cplx s = 0.0;
      for (int i = 0; i < blk; i++) s += Ar[i + j * blk] * U[i * n + k];
      E[j * n + k] = s;
    }
  free(U); free(Ar); free(Mr);
  return 0;
}

static void residuals(og_mg *mg, int blk, const cplx *E, const double *w,
                      double *relres, double *absres)
{
  const og_grid *g = &mg->g[mg->L];
  const size_t n = edim(mg, mg->L);
  cplx *r = (cplx *)malloc(sizeof(cplx) * n);
  cplx *m = (cplx *)malloc(sizeof(cplx) * n);
  for (int i = 0; i < blk; i++) {
    memset(r, 0, sizeof(cplx) * n);
    og_addeval(g, 1.0, 0.0, 1.0 - w[i], 0.0, (const double *)(E + i * n), (double *)r, mg->quirk);
    memset(m, 0, sizeof(cplx) * n);
    og_addeval(g, 0.0, 0.0, 1.0, 0.0, (const double *)(E + i * n), (double *)m, mg->quirk);
    double rn = og_norm2((int)n, (const double *)r);
    double mn = og_norm2((int)n, (const double *)m);
    absres[i] = rn;
    relres[i] = rn / (fabs(w[i] - 1.0) * mn);
  }
  free(r); free(m);
}

int og_block_solve(og_mg *mg, int neig, int blk, double tol, int maxit,
                   int prec_steps, int gradient_steps, double *Ed,
                   double *lam_out, double *relres, double *absres)
{
  const og_grid *g = &mg->g[mg->L];
  const size_t n = edim(mg, mg->L);
  cplx *E = (cplx *)Ed;
  double *w = (double *)malloc(sizeof(double) * blk);
  for (int i = 0; i < blk; i++) {
    if (is_gamma(g)) og_center(g, (double *)(E + i * n));
    og_lift(mg, gradient_steps, (double *)(E + i * n));
  }
  int it = 0;
  if (ritz(mg, blk, E, w) == 0) {
    residuals(mg, blk, E, w, relres, absres);
    for (;;) {
      double worst = 0.0;
      for (int i = 0; i < neig; i++) if (!(relres[i] <= worst)) worst = relres[i];
      if (worst <= tol || it >= maxit) break;
      for (int i = 0; i < blk; i++)
        og_pinvit_vector(mg, prec_steps, gradient_steps, w[i], (double *)(E + i * n));
      if (ritz(mg, blk, E, w) != 0) break;
      residuals(mg, blk, E, w, relres, absres);
      it++;
    }
  }
  for (int i = 0; i < blk; i++) lam_out[i] = w[i] - 1.0;
  free(w);
  return it;
}

void og_eta_gradient(const og_grid *g, const double *ed, double *grad)
{
  const int nx = g->nx, ny = g->ny;
  const cplx *e = (const cplx *)ed;
  double m[2];
  og_massprod(g, ed, ed, m);
  const double scale = 1.0 / (g->hx * g->hy) / m[0];
#pragma omp parallel for
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      cplx c = curl(g, e, i, j);
      grad[i + j * nx] = creal(c * conj(c)) * scale;
    }
}

/* ---- smooth test fields for the interpolation and quadrature routines of
   the module API (edge2d.c:643-1206, node2d.c:93-385), in the callback
   signatures those routines take; the same formulas are restated in numpy
   by the Python bindings (field_edge, field_node). */
void og_field_edge(const double *p, cplx *f, void *data)
{
  (void)data;
  const double x = p[0], y = p[1], tau = 6.283185307179586;
  f[0] = cos(tau * y) + 0.2 * I * x + 0.1 * x * y;
  f[1] = sin(tau * x) * (1.0 + 0.5 * I) + y;
}

cplx og_field_node(const double *p, void *data)
{
  (void)data;
  const double x = p[0], y = p[1], tau = 6.283185307179586;
  return (1.0 + 0.5 * x) * cexp(I * tau * (x + 2.0 * y)) + 0.3 * I * y * y;
}
