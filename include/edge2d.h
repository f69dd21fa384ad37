/*
 * edge2d.h -- lowest-order Nedelec edge functions on a periodic grid
 * (reference edge2d.h:12-293).  Signatures and meaning follow the reference;
 * the coefficient arrays live in GPU memory (x: nx*ny x-edge values, y: the
 * nx*ny y-edge values right after them), so callers must not dereference
 * x or y -- the reference's drivers never do.  Vector functions are
 * collectives of the calling OpenMP team, as in the reference: call them from
 * all threads of a parallel region or serially.
 */
#ifndef EDGE2D_H
#define EDGE2D_H

#include <stdbool.h>
#include <stdlib.h>

#include "grid2d.h"

typedef struct {
  const grid2d *gr;
  field *x;   /* device: x-edge (i,j) at x[i + j*nx] */
  field *y;   /* device: y-edge (i,j) at y[i + j*nx] (y == x + nx*ny) */
} edge2d;

#include "linalg.h"
#include "node2d.h"

/* ---- constructor, destructor, utilities (edge2d.c:22-191) */
edge2d *new_edge2d(const grid2d *gr);
void del_edge2d(edge2d *x);
size_t getdimension_edge2d(const edge2d *x);
size_t getsize_edge2d(const edge2d *x);
void zero_edge2d(edge2d *x);
void copy_edge2d(const edge2d *x, edge2d *y);
void swap_edge2d(edge2d *x, edge2d *y);
void random_edge2d(edge2d *x);

/* ---- BLAS-1 (edge2d.c:806-920, 1207-1333) */
void scale_edge2d(field alpha, edge2d *x);
void add_edge2d(field alpha, const edge2d *x, edge2d *y);
real norm2_edge2d(const edge2d *x);
real normmax_edge2d(const edge2d *x);
field dotprod_edge2d(const edge2d *x, const edge2d *y);
void nullprod_edge2d(const edge2d *x, field *xprod, field *yprod);
void nulladd_edge2d(field xalpha, field yalpha, edge2d *x);
void center_edge2d(edge2d *x);

/* ---- operators (edge2d.c:1335-2027) */
void addeval_edge2d(field alpha, field beta, const edge2d *x, edge2d *y);
field energyprod_edge2d(field alpha, field beta, const edge2d *x, const edge2d *y);
field massprod_edge2d(const edge2d *x, const edge2d *y);

/* ---- AFW vertex-patch Gauss-Seidel (edge2d.c:2656-3042); four-colour order */
void gsforward_edge2d(field alpha, field beta, const edge2d *b, edge2d *x);
void gsbackward_edge2d(field alpha, field beta, const edge2d *b, edge2d *x);
void gssymm_edge2d(field alpha, field beta, const edge2d *b, edge2d *x);

/* ---- grid transfer (edge2d.c:5583-5705) */
void prolongation_edge2d(field alpha, const edge2d *c, edge2d *f);
void restriction_edge2d(field alpha, const edge2d *f, edge2d *c);

/* ---- multigrid (edge2d.c:5707-5754) */
void vcycle_edge2d(int L, int nu, const matrix *Ac, edge2d **b, edge2d **x, edge2d **d);

/* ---- dense (edge2d.c:5870-6152) */
void densematrix_edge2d(field alpha, field beta, const grid2d *gr, matrix *a);
void denseto_edge2d(const field *x, edge2d *y);
void densefrom_edge2d(const edge2d *x, field *y);

/* ---- Householder M-orthonormalisation (edge2d.c:6154-6296) */
void unitindices_edge2d(int idx, const edge2d *x, int *i, int *j);
void unit_edge2d(int idx, edge2d *x);
field massprod_xunit_edge2d(int i, int j, const edge2d *y);
void add_xunit_edge2d(field alpha, int i, int j, edge2d *y);
void buildhouseholder_edge2d(int idx, edge2d *x, field *tau);
void applyhouseholder_edge2d(const edge2d *v, field tau, edge2d *x);
void orthonormalize_edge2d(int k, edge2d **x, field *tau, edge2d *v);

/* ---- PINVIT vector step with gradient elimination (edge2d.c:6298-6328) */
void pinvit_edge2d(int l, int smoother_steps, int prec_steps, int gradient_steps, real lambda, edge2d *e,
                   const matrix *Ae, edge2d **b, edge2d **x, edge2d **d, const matrix *An, node2d **bg,
                   node2d **xg, node2d **dg);

#endif
