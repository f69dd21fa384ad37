/* Host dense linear algebra and coarse-grid assembly (internal). */
#ifndef MX_DENSE_H
#define MX_DENSE_H

#include <complex.h>

/* Lower Cholesky in place; 0 or the 1-based index of a non-positive pivot. */
int mxd_potrf(int n, double complex *a);
/* Solve L L^H x = b in place. */
void mxd_potrs(int n, const double complex *l, double complex *b);
/* Hermitian eigenproblem from the lower triangle; a <- eigenvectors. */
int mxd_heev(int n, double complex *a, double *w);
/* A x = w B x (lower triangles); a <- B-orthonormal eigenvectors,
   b <- Cholesky factor.  0, n + k for a non-definite B, or -1 (memory). */
int mxd_hegv(int n, double complex *a, double complex *b, double *w);
/* mxd_hegv that also reports B as not definite when a Cholesky pivot of B
   falls to rtol times its diagonal entry or below (numerically dependent
   columns, e.g. repeated vectors in a Gram matrix). */
int mxd_hegv_rel(int n, double complex *a, double complex *b, double *w, double rtol);

void mxd_dense_edge(int nx, int ny, double hx, double hy, double complex xf, double complex yf,
                    const double *eta, double complex alpha, double complex beta, double complex *a, int lda);
void mxd_dense_node(int nx, int ny, double hx, double hy, double complex xf, double complex yf, int fixnull,
                    double complex *a, int lda);

#endif
