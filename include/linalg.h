/*
 * linalg.h -- small dense matrices on the host (reference linalg.h:7-31).
 * Column-major, ld = rows.  The factorisations are this library's own
 * routines (no LAPACK); return codes follow LAPACK's info convention.
 */
#ifndef LINALG_H
#define LINALG_H

#include "settings.h"

typedef struct {
  int rows;
  int cols;
  int ld;
  field *a;
} matrix;

matrix *new_matrix(int rows, int cols);
void del_matrix(matrix *a);

/* Lower Cholesky factor in place (zpotrf "L", linalg.c:53-63). */
int potrf_matrix(matrix *a);
/* b <- (L L^H)^{-1} b with the factor from potrf_matrix (linalg.c:98-113). */
void potrsv_matrix(const matrix *a, field *b);
/* Hermitian eigenproblem, lower triangle, vectors in a (zheev, linalg.c:254-283). */
int syev_matrix(matrix *a, real *lambda);
/* A x = lambda B x, lower triangles, B-orthonormal vectors in a
   (zhegv itype 1, linalg.c:285-317). */
int sygv_matrix(matrix *a, matrix *b, real *lambda);

#endif
