/*
 * node2d.h -- bilinear nodal functions on a periodic grid
 * (reference node2d.h:9-158).  v lives in GPU memory (nx*ny values,
 * node (i,j) at v[i + j*nx]); callers must not dereference it.
 */
#ifndef NODE2D_H
#define NODE2D_H

#include <stdbool.h>

#include "grid2d.h"

typedef struct {
  const grid2d *gr;
  field *v;   /* device */
} node2d;

#include "edge2d.h"
#include "linalg.h"

node2d *new_node2d(const grid2d *gr);
void del_node2d(node2d *x);
void zero_node2d(node2d *x);
void copy_node2d(const node2d *x, node2d *y);
void random_node2d(node2d *x);

/* given functions (node2d.c:93-385): interpolation at the nodes, L2
   functional (buf: scratch) and L2 norm of x - func (of x when func is null)
   with the 2 x 2 Gauss rule per cell */
void interpolate_node2d(node2d *x, field (*func)(const real *x, void *data), void *data);
void l2functional_node2d(node2d *x, node2d *buf, field (*func)(const real *x, void *data), void *data);
real l2norm_node2d(const node2d *x, field (*func)(const real *x, void *data), void *data);

void scale_node2d(field alpha, node2d *x);
void add_node2d(field alpha, const node2d *x, node2d *y);
real norm2_node2d(const node2d *x);
real normmax_node2d(const node2d *x);
field dotprod_node2d(const node2d *x, const node2d *y);
void center_node2d(node2d *x);

/* y += alpha A x, A the Q1 nodal Laplacian (node2d.c:516-627) */
void addeval_node2d(field alpha, const node2d *x, node2d *y);
/* pointwise Gauss-Seidel (node2d.c:629-1138); four-colour order */
void gsforward_node2d(field alpha, const node2d *b, node2d *x);
void gsbackward_node2d(field alpha, const node2d *b, node2d *x);
void gssymm_node2d(field alpha, const node2d *b, node2d *x);

void prolongation_node2d(field alpha, const node2d *c, node2d *f);
void restriction_node2d(field alpha, const node2d *f, node2d *c);
/* y += alpha G x and y += alpha G^H x (node2d.c:1275-1350) */
void gradient_node2d(field alpha, const node2d *x, edge2d *y);
void adjgradient_node2d(field alpha, const edge2d *x, node2d *y);

void vcycle_node2d(int L, int nu, const matrix *Ac, node2d **b, node2d **x, node2d **d);
/* V-cycle preconditioned conjugate gradients for the Q1 Laplacian
   (node2d.c:1392-1427) */
void mgcginit_node2d(int L, int nu, const node2d *b, node2d *x, node2d *p, node2d *a, node2d **r, node2d **q,
                     node2d **d);
void mgcgstep_node2d(int L, int nu, const node2d *b, node2d *x, node2d *p, node2d *a, node2d **r, node2d **q,
                     node2d **d);

void densematrix_node2d(bool fixnull, const grid2d *gr, matrix *a);
void denseto_node2d(const field *x, node2d *y);
void densefrom_node2d(const node2d *x, field *y);

#endif
