/*
 * dense.c -- small dense complex linear algebra on the host: the operations
 * the reference takes from LAPACK (linalg.c:53-63 zpotrf, 98-113 ztrsm x2,
 * 260-283 zheev, 291-317 zhegv itype 1).  Sizes here are at most a few
 * hundred (coarse grids, Ritz blocks), so plain loops suffice.
 *
 * Matrices are column-major double complex with leading dimension n.
 */
#include "mx_dense.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A_(i, j) a[(i) + (size_t)(j) * n]

/* Cholesky that also fails when a pivot falls to rtol times its column's
   diagonal entry or below (rtol = 0: only non-positive pivots fail). */
static int potrf_rel(int n, double complex *a, double rtol)
{
  for (int c = 0; c < n; c++) {
    const double d0 = creal(A_(c, c));
    double s = d0;
    for (int k = 0; k < c; k++) s -= creal(A_(c, k) * conj(A_(c, k)));
    if (!(s > 0.0) || !(s > rtol * d0)) return c + 1;
    double d = sqrt(s);
    A_(c, c) = d;
    for (int r = c + 1; r < n; r++) {
      double complex t = A_(r, c);
      for (int k = 0; k < c; k++) t -= A_(r, k) * conj(A_(c, k));
      A_(r, c) = t / d;
    }
    for (int r = 0; r < c; r++) A_(r, c) = 0.0;
  }
  return 0;
}

int mxd_potrf(int n, double complex *a) { return potrf_rel(n, a, 0.0); }

void mxd_potrs(int n, const double complex *a, double complex *b)
{
  for (int r = 0; r < n; r++) {
    double complex t = b[r];
    for (int k = 0; k < r; k++) t -= A_(r, k) * b[k];
    b[r] = t / A_(r, r);
  }
  for (int r = n - 1; r >= 0; r--) {
    double complex t = b[r];
    for (int k = r + 1; k < n; k++) t -= conj(A_(k, r)) * b[k];
    b[r] = t / conj(A_(r, r));
  }
}

/* Cyclic Jacobi on a full Hermitian matrix (destroyed); w ascending, V
   eigenvectors in columns. */
static void heev_jacobi(int n, double complex *a, double *w, double complex *v)
{
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) v[i + (size_t)j * n] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        double t = creal(A_(i, j) * conj(A_(i, j)));
        if (i == j) dia += t;
        else off += t;
      }
    if (off == 0.0 || off <= 1e-32 * dia) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double complex z = A_(p, q);
        double r = cabs(z);
        if (r == 0.0) continue;
        double app = creal(A_(p, p)), aqq = creal(A_(q, q));
        double tau = (aqq - app) / (2.0 * r);
        double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        double complex e = conj(z) / r;
        for (int k = 0; k < n; k++) {
          double complex kp = A_(k, p), kq = A_(k, q);
          A_(k, p) = cs * kp - sn * e * kq;
          A_(k, q) = sn * kp + cs * e * kq;
        }
        for (int k = 0; k < n; k++) {
          double complex pk = A_(p, k), qk = A_(q, k);
          A_(p, k) = cs * pk - sn * conj(e) * qk;
          A_(q, k) = sn * pk + cs * conj(e) * qk;
        }
        A_(p, q) = 0.0;
        A_(q, p) = 0.0;
        A_(p, p) = creal(A_(p, p));
        A_(q, q) = creal(A_(q, q));
        for (int k = 0; k < n; k++) {
          double complex kp = v[k + (size_t)p * n], kq = v[k + (size_t)q * n];
          v[k + (size_t)p * n] = cs * kp - sn * e * kq;
          v[k + (size_t)q * n] = sn * kp + cs * e * kq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = creal(A_(i, i));
  for (int i = 0; i < n; i++) {
    int m = i;
    for (int k = i + 1; k < n; k++)
      if (w[k] < w[m]) m = k;
    if (m != i) {
      double t = w[i];
      w[i] = w[m];
      w[m] = t;
      for (int k = 0; k < n; k++) {
        double complex s = v[k + (size_t)i * n];
        v[k + (size_t)i * n] = v[k + (size_t)m * n];
        v[k + (size_t)m * n] = s;
      }
    }
  }
}

int mxd_heev(int n, double complex *a, double *w)
{
  if (n == 0) return 0;
  double complex *c = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  double complex *v = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  if (!c || !v) { free(c); free(v); return -1; }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      c[i + (size_t)j * n] = (i >= j) ? A_(i, j) : conj(A_(j, i));
  for (int i = 0; i < n; i++) c[i + (size_t)i * n] = creal(c[i + (size_t)i * n]);
  heev_jacobi(n, c, w, v);
  memcpy(a, v, sizeof(double complex) * (size_t)n * n);
  free(v);
  free(c);
  return 0;
}

int mxd_hegv(int n, double complex *a, double complex *b, double *w) { return mxd_hegv_rel(n, a, b, w, 0.0); }

int mxd_hegv_rel(int n, double complex *a, double complex *b, double *w, double rtol)
{
  if (n == 0) return 0;
  int info = potrf_rel(n, b, rtol);
  if (info) return n + info;
  double complex *c = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  double complex *v = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  if (!c || !v) { free(c); free(v); return -1; }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      c[i + (size_t)j * n] = (i >= j) ? A_(i, j) : conj(A_(j, i));
  for (int i = 0; i < n; i++) c[i + (size_t)i * n] = creal(c[i + (size_t)i * n]);
  /* c <- L^{-1} c L^{-H} */
#define B_(i, j) b[(i) + (size_t)(j) * n]
#define C_(i, j) c[(i) + (size_t)(j) * n]
  for (int j = 0; j < n; j++)
    for (int r = 0; r < n; r++) {
      double complex t = C_(r, j);
      for (int k = 0; k < r; k++) t -= B_(r, k) * C_(k, j);
      C_(r, j) = t / B_(r, r);
    }
  for (int i = 0; i < n; i++)
    for (int r = 0; r < n; r++) {
      double complex t = C_(i, r);
      for (int k = 0; k < r; k++) t -= C_(i, k) * conj(B_(r, k));
      C_(i, r) = t / conj(B_(r, r));
    }
  /* symmetrise against rounding */
  for (int j = 0; j < n; j++) {
    C_(j, j) = creal(C_(j, j));
    for (int i = j + 1; i < n; i++) {
      double complex s = 0.5 * (C_(i, j) + conj(C_(j, i)));
      C_(i, j) = s;
      C_(j, i) = conj(s);
    }
  }
  heev_jacobi(n, c, w, v);
  /* x = L^{-H} y */
  for (int j = 0; j < n; j++)
    for (int r = n - 1; r >= 0; r--) {
      double complex t = v[r + (size_t)j * n];
      for (int k = r + 1; k < n; k++) t -= conj(B_(k, r)) * v[k + (size_t)j * n];
      v[r + (size_t)j * n] = t / conj(B_(r, r));
    }
#undef B_
#undef C_
  memcpy(a, v, sizeof(double complex) * (size_t)n * n);
  free(v);
  free(c);
  return 0;
}

/* ---------------------------------------------------------------- assembly */

static double complex wrap_phase(int nx, int ny, double complex xf, double complex yf, int *i, int *j)
{
  double complex ph = 1.0;
  if (*i < 0) { *i += nx; ph *= conj(xf); }
  else if (*i >= nx) { *i -= nx; ph *= xf; }
  if (*j < 0) { *j += ny; ph *= conj(yf); }
  else if (*j >= ny) { *j -= ny; ph *= yf; }
  return ph;
}

/* alpha K + beta M (edge2d.c:5870-6110), assembled from the cell matrices. */
void mxd_dense_edge(int nx, int ny, double hx, double hy, double complex xf, double complex yf,
                    const double *eta, double complex alpha, double complex beta, double complex *a, int lda)
{
  const int N = nx * ny;
  const double complex axy = alpha / hx / hy;
  const double complex mx = beta * hy / hx / 6.0, my = beta * hx / hy / 6.0;
  for (int j = 0; j < 2 * N; j++)
    for (int i = 0; i < 2 * N; i++) a[i + (size_t)j * lda] = 0.0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int gi[4];
      double complex ph[4];
      int di[4] = {0, 0, 0, 1}, dj[4] = {0, 1, 0, 0}, off[4] = {0, 0, N, N};
      for (int p = 0; p < 4; p++) {
        int ii = i + di[p], jj = j + dj[p];
        ph[p] = wrap_phase(nx, ny, xf, yf, &ii, &jj);
        gi[p] = off[p] + ii + jj * nx;
      }
      const double sgn[4] = {1.0, -1.0, -1.0, 1.0};
      double complex ke[4][4];
      for (int p = 0; p < 4; p++)
        for (int q = 0; q < 4; q++) ke[p][q] = axy * eta[i + j * nx] * sgn[p] * sgn[q];
      ke[0][0] += 2.0 * mx; ke[1][1] += 2.0 * mx; ke[0][1] += mx; ke[1][0] += mx;
      ke[2][2] += 2.0 * my; ke[3][3] += 2.0 * my; ke[2][3] += my; ke[3][2] += my;
      for (int p = 0; p < 4; p++)
        for (int q = 0; q < 4; q++) a[gi[p] + (size_t)gi[q] * lda] += conj(ph[p]) * ke[p][q] * ph[q];
    }
}

/* Q1 nodal Laplacian (node2d.c:1429-1602) with the Gamma-point fill of
   node2d.c:1454-1458 when fixnull is set. */
void mxd_dense_node(int nx, int ny, double hx, double hy, double complex xf, double complex yf, int fixnull,
                    double complex *a, int lda)
{
  const int n = nx * ny;
  double fill = 0.0;
  if (fixnull && cabs(xf - 1.0) < 1e-8 && cabs(yf - 1.0) < 1e-8) fill = 0.1 / ((double)n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) a[i + (size_t)j * lda] = fill;
  const double rx = hy / hx, ry = hx / hy;
  const double S[2][2] = {{1.0, -1.0}, {-1.0, 1.0}};
  const double T[2][2] = {{1.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 3.0}};
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int gi[4];
      double complex ph[4];
      for (int p = 0; p < 4; p++) {
        int ii = i + (p & 1), jj = j + (p >> 1);
        ph[p] = wrap_phase(nx, ny, xf, yf, &ii, &jj);
        gi[p] = ii + jj * nx;
      }
      for (int p = 0; p < 4; p++)
        for (int q = 0; q < 4; q++) {
          int px = p & 1, py = p >> 1, qx = q & 1, qy = q >> 1;
          double ke = rx * S[px][qx] * T[py][qy] + ry * T[px][qx] * S[py][qy];
          a[gi[p] + (size_t)gi[q] * lda] += conj(ph[p]) * ke * ph[q];
        }
    }
}
