/*
 * grid2d.h -- periodic grid and permittivity (reference grid2d.h:7-78).
 *
 * Same fields and meaning as the reference; `dev` (appended) is the device
 * mirror of eps.  new_grid2d, coarsen_grid2d and setpattern_grid2d keep the
 * mirror current; code that writes gr->eps[] directly calls sync_grid2d.
 * xfactor/yfactor are read from the struct on every call, so callers may set
 * them directly as exp_bloch2d.c:794-797 does.
 */
#ifndef GRID2D_H
#define GRID2D_H

#include "settings.h"

typedef struct {
  real hx;          /* mesh width in x */
  real hy;          /* mesh width in y */
  int nx;           /* intervals in x */
  int ny;           /* intervals in y */
  real xbloch;      /* Bloch parameter kx */
  real ybloch;      /* Bloch parameter ky */
  field xfactor;    /* e^{i kx} */
  field yfactor;    /* e^{i ky} */
  real *eps;        /* 1/eps per cell, eps[i + j*nx] (host) */
  void *dev;        /* device mirror (library-private) */
} grid2d;

grid2d *new_grid2d(int nx, int ny, real hx, real hy);
void del_grid2d(grid2d *gr);
grid2d *coarsen_grid2d(const grid2d *fg);
/* Writes a binary PGM of eps; the reference draws with cairo (grid2d.c:70-142). */
void draw2d_grid2d(const grid2d *gr, const char *filename);
/* Upload gr->eps to the device after direct writes. */
void sync_grid2d(grid2d *gr);

typedef struct {
  real *rectangle;  /* (eps, x0, y0, width, height) per rectangle */
  int rectangles;
  real *circle;     /* (eps, cx, cy, radius) per circle */
  int circles;
  real eps_base;    /* permittivity of the background */
} epspattern;

epspattern *new_epspattern(real eps_base, int rectangles, int circles);
void del_epspattern(epspattern *pat);
void setpattern_grid2d(const epspattern *pat, grid2d *gr);

#endif
