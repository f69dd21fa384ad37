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

/* ---- given functions (edge2d.c:643-1206): interpolation, L2 functional
   (xbuf, ybuf: scratch that receives the function's components at the edge
   midpoints) and the L2 norm of x - func (of x when func is null).  func
   writes the field's two components at the point x. */
void interpolate_edge2d(edge2d *x, void (*func)(const real *x, field *fx, void *data), void *data);
void l2functional_edge2d(edge2d *x, edge2d *xbuf, edge2d *ybuf,
                         void (*func)(const real *x, field *fx, void *data), void *data);
real l2norm_edge2d(const edge2d *x, void (*func)(const real *x, field *fx, void *data), void *data);

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

/* ---- pointwise (one edge at a time) Gauss-Seidel (edge2d.c:3892-5600), the
   smoother of the Hiptmair cycle; four-colour order */
void gsforward_simple_edge2d(field alpha, field beta, const edge2d *b, edge2d *x);
void gsbackward_simple_edge2d(field alpha, field beta, const edge2d *b, edge2d *x);
void gssymm_simple_edge2d(field alpha, field beta, const edge2d *b, edge2d *x);

/* ---- grid transfer (edge2d.c:5583-5705) */
void prolongation_edge2d(field alpha, const edge2d *c, edge2d *f);
void restriction_edge2d(field alpha, const edge2d *f, edge2d *c);

/* ---- multigrid (edge2d.c:5707-5868): AFW V-cycle, Hiptmair V-cycle,
   V-cycle preconditioned conjugate gradients */
void vcycle_edge2d(int L, int nu, const matrix *Ac, edge2d **b, edge2d **x, edge2d **d);
void hcycle_edge2d(int L, int nu, const matrix *Ac, edge2d **b, edge2d **x, edge2d **d, node2d **bg,
                   node2d **xg);
void mgcginit_edge2d(int L, int nu, const edge2d *b, edge2d *x, edge2d *p, edge2d *a, edge2d **r, edge2d **q,
                     edge2d **d);
void mgcgstep_edge2d(int L, int nu, const edge2d *b, edge2d *x, edge2d *p, edge2d *a, edge2d **r, edge2d **q,
                     edge2d **d);

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

/* ---- eigenvector files (edge2d.h:35-64, 105-116; edge2d.c:385-640).
   The same dimensions, variables and values as the reference writes through
   libnetcdf, written here without it as a netCDF 64-bit-offset file.  The
   reference's ncfile fields come first; the rest belongs to this writer. */
typedef struct {
  int file;
  int nc_blochsteps;
  int blochsteps;
  int nc_eigenvalues;
  int eigenvalues;
  int nc_graphx;
  int graphx;
  int nc_graphy;
  int graphy;
  int nc_lambda;
  int nc_vectorxr;
  int nc_vectorxi;
  int nc_vectoryr;
  int nc_vectoryi;
  int nc_xbloch;
  int nc_ybloch;
  int nc_xfactorr;
  int nc_xfactori;
  int nc_yfactorr;
  int nc_yfactori;
  long long begin[16]; /* file offset of each variable's data */
} ncfile;

ncfile *new_ncfile(const edge2d *x, int eigs, int blochsteps, int gx, int gy, const epspattern *pat,
                   const char *filename);
void del_ncfile(ncfile *nf);
void write_ncfile(ncfile *nf, int blochstep, const real *lambda, const edge2d **eigs);

/* ---- picture of a field (edge2d.h:99-103).  The reference draws arrows with
   cairo into PNG or PDF; this writes the field magnitude, sampled on a
   gx x gy grid as write_ncfile samples it, as a binary PGM (whatever the
   file name's extension). */
void cairodraw_edge2d(const edge2d *x, int gx, int gy, const epspattern *pat, const char *filename);

#endif
