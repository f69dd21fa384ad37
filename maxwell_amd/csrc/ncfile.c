/*
 * ncfile.c -- the Bloch drivers' eigenvector files (edge2d.c:385-640) and
 * field pictures (edge2d.c:195-383), without libnetcdf or cairo.
 *
 * new_ncfile defines what the reference defines through libnetcdf: the
 * dimensions blochsteps, eigenvalues, graphx, graphy, rectangles (5 per
 * rectangle of the pattern), circles (4 per circle), and the double variables
 * xbloch, ybloch, xfactorr, xfactori, yfactorr, yfactori (per Bloch step),
 * rectangle and circle (the pattern, written at once), lambda (step x
 * eigenvalue: lambda - 1, the eigenvalue of K), vectorxr/xi/yr/yi (step x
 * eigenvalue x graphx x graphy: the field sampled on a gx x gy grid).  A
 * dimension of length 0 is the unlimited one, as NC_UNLIMITED is 0 for
 * nc_def_dim.
 *
 * The file is the netCDF 64-bit-offset format ("CDF" 2): big-endian header
 * of dimension and variable lists, then the fixed-size variables in the
 * order of definition, prefilled with the default double fill value as
 * libnetcdf's fill mode does, then the record variables (here: none or one
 * with zero records).  The reference asks libnetcdf for the CDF-5 variant
 * (NC_64BIT_DATA); both are read by every netCDF library, hold the same
 * dimensions, variables and values, and differ only in the widths of the
 * header's counts.
 */
#include <fcntl.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "edge2d.h"
#include "grid2d.h"

#define TAG_DIMENSION 0x0Au
#define TAG_VARIABLE 0x0Bu
#define TYPE_DOUBLE 6u
#define FILL_DOUBLE 9.9692099683868690e+36 /* libnetcdf's NC_FILL_DOUBLE */

static void nc_fail(const char *what)
{
  fprintf(stderr, "maxwell_amd: netCDF file: %s\n", what);
  abort();
}

typedef struct {
  unsigned char *p;
  size_t n, cap;
} hbuf;

static void put_bytes(hbuf *b, const void *src, size_t n)
{
  if (b->n + n > b->cap) {
    b->cap = 2 * (b->n + n) + 256;
    b->p = (unsigned char *)realloc(b->p, b->cap);
    if (!b->p) nc_fail("out of memory");
  }
  memcpy(b->p + b->n, src, n);
  b->n += n;
}

static void put32(hbuf *b, uint32_t v)
{
  unsigned char c[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8), (unsigned char)v};
  put_bytes(b, c, 4);
}

static void put64(hbuf *b, uint64_t v)
{
  put32(b, (uint32_t)(v >> 32));
  put32(b, (uint32_t)v);
}

static void put_name(hbuf *b, const char *s)
{
  static const unsigned char zero[4] = {0, 0, 0, 0};
  const size_t n = strlen(s);
  put32(b, (uint32_t)n);
  put_bytes(b, s, n);
  put_bytes(b, zero, (4 - n % 4) % 4);
}

static uint64_t be64(double v)
{
  uint64_t u;
  memcpy(&u, &v, 8);
  return __builtin_bswap64(u);
}

static void put_doubles(int fd, long long off, const double *v, size_t n)
{
  uint64_t tmp[512];
  while (n) {
    const size_t k = n < 512 ? n : 512;
    for (size_t i = 0; i < k; i++) tmp[i] = be64(v[i]);
    const ssize_t w = pwrite(fd, tmp, 8 * k, (off_t)off);
    if (w != (ssize_t)(8 * k)) nc_fail("write failed");
    off += 8 * (long long)k;
    v += k;
    n -= k;
  }
}

enum { D_STEPS, D_EIGS, D_GX, D_GY, D_RECT, D_CIRC, NDIM };

ncfile *new_ncfile(const edge2d *x, int eigs, int blochsteps, int gx, int gy, const epspattern *pat,
                   const char *filename)
{
  (void)x; /* the reference does not read it either */
  if (eigs < 1 || blochsteps < 1 || gx < 1 || gy < 1) nc_fail("dimensions must be positive");
  const char *dname[NDIM] = {"blochsteps", "eigenvalues", "graphx", "graphy", "rectangles", "circles"};
  long long dlen[NDIM] = {blochsteps, eigs, gx, gy, pat ? 5LL * pat->rectangles : 0, pat ? 4LL * pat->circles : 0};
  int recdim = -1;
  for (int d = 0; d < NDIM; d++)
    if (dlen[d] == 0) {
      if (recdim >= 0) nc_fail("more than one dimension of length 0 (unlimited)");
      recdim = d;
    }

  ncfile *nf = (ncfile *)calloc(1, sizeof(ncfile));
  if (!nf) nc_fail("out of memory");
  nf->nc_blochsteps = D_STEPS;
  nf->blochsteps = blochsteps;
  nf->nc_eigenvalues = D_EIGS;
  nf->eigenvalues = eigs;
  nf->nc_graphx = D_GX;
  nf->graphx = gx;
  nf->nc_graphy = D_GY;
  nf->graphy = gy;

  /* variables in the reference's order of definition */
  struct {
    const char *name;
    int ndim, dim[4];
  } var[13];
  int nv = 0, *id[13];
  const char *scal[6] = {"xbloch", "ybloch", "xfactorr", "xfactori", "yfactorr", "yfactori"};
  int *scal_id[6] = {&nf->nc_xbloch, &nf->nc_ybloch, &nf->nc_xfactorr, &nf->nc_xfactori, &nf->nc_yfactorr,
                     &nf->nc_yfactori};
  for (int k = 0; k < 6; k++) {
    var[nv].name = scal[k];
    var[nv].ndim = 1;
    var[nv].dim[0] = D_STEPS;
    id[nv++] = scal_id[k];
  }
  int v_rect = -1, v_circ = -1;
  if (pat) {
    v_rect = nv;
    var[nv].name = "rectangle";
    var[nv].ndim = 1;
    var[nv].dim[0] = D_RECT;
    id[nv++] = NULL;
    v_circ = nv;
    var[nv].name = "circle";
    var[nv].ndim = 1;
    var[nv].dim[0] = D_CIRC;
    id[nv++] = NULL;
  }
  var[nv].name = "lambda";
  var[nv].ndim = 2;
  var[nv].dim[0] = D_STEPS;
  var[nv].dim[1] = D_EIGS;
  id[nv++] = &nf->nc_lambda;
  const char *vec[4] = {"vectorxr", "vectorxi", "vectoryr", "vectoryi"};
  int *vec_id[4] = {&nf->nc_vectorxr, &nf->nc_vectorxi, &nf->nc_vectoryr, &nf->nc_vectoryi};
  for (int k = 0; k < 4; k++) {
    var[nv].name = vec[k];
    var[nv].ndim = 4;
    var[nv].dim[0] = D_STEPS;
    var[nv].dim[1] = D_EIGS;
    var[nv].dim[2] = D_GX;
    var[nv].dim[3] = D_GY;
    id[nv++] = vec_id[k];
  }
  for (int v = 0; v < nv; v++)
    if (id[v]) *id[v] = v;

  /* bytes per variable (per record for a record variable) */
  long long vsize[13];
  int isrec[13];
  for (int v = 0; v < nv; v++) {
    long long s = 8;
    isrec[v] = var[v].dim[0] == recdim;
    for (int d = isrec[v] ? 1 : 0; d < var[v].ndim; d++) s *= dlen[var[v].dim[d]];
    vsize[v] = s;
  }

  /* header, with the data offsets patched in after its length is known */
  hbuf h = {0};
  for (int pass = 0; pass < 2; pass++) {
    h.n = 0;
    put_bytes(&h, "CDF\002", 4);
    put32(&h, 0); /* numrecs */
    put32(&h, TAG_DIMENSION);
    put32(&h, NDIM);
    for (int d = 0; d < NDIM; d++) {
      put_name(&h, dname[d]);
      put32(&h, (uint32_t)dlen[d]);
    }
    put32(&h, 0); /* no global attributes */
    put32(&h, 0);
    put32(&h, TAG_VARIABLE);
    put32(&h, (uint32_t)nv);
    for (int v = 0; v < nv; v++) {
      put_name(&h, var[v].name);
      put32(&h, (uint32_t)var[v].ndim);
      for (int d = 0; d < var[v].ndim; d++) put32(&h, (uint32_t)var[v].dim[d]);
      put32(&h, 0); /* no variable attributes */
      put32(&h, 0);
      put32(&h, TYPE_DOUBLE);
      put32(&h, vsize[v] < 0xFFFFFFFFLL ? (uint32_t)vsize[v] : 0xFFFFFFFFu);
      put64(&h, (uint64_t)(pass ? nf->begin[v] : 0));
    }
    if (pass == 0) {
      long long o = (long long)h.n;
      for (int v = 0; v < nv; v++)
        if (!isrec[v]) nf->begin[v] = o, o += vsize[v];
      for (int v = 0; v < nv; v++)
        if (isrec[v]) nf->begin[v] = o, o += vsize[v];
    }
  }

  const int fd = open(filename, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) nc_fail("cannot create the file");
  nf->file = fd;
  if (pwrite(fd, h.p, h.n, 0) != (ssize_t)h.n) nc_fail("write failed");
  free(h.p);

  /* fill mode: every fixed-size variable starts as the fill value */
  double fillbuf[4096];
  for (int k = 0; k < 4096; k++) fillbuf[k] = FILL_DOUBLE;
  for (int v = 0; v < nv; v++) {
    if (isrec[v]) continue;
    long long left = vsize[v] / 8, off = nf->begin[v];
    while (left > 0) {
      const long long k = left < 4096 ? left : 4096;
      put_doubles(fd, off, fillbuf, (size_t)k);
      off += 8 * k;
      left -= k;
    }
  }
  if (pat) {
    if (!isrec[v_rect]) put_doubles(fd, nf->begin[v_rect], pat->rectangle, 5 * (size_t)pat->rectangles);
    if (!isrec[v_circ]) put_doubles(fd, nf->begin[v_circ], pat->circle, 4 * (size_t)pat->circles);
  }
  return nf;
}

void del_ncfile(ncfile *nf)
{
  if (close(nf->file)) nc_fail("close failed");
  free(nf);
}

/* The field of one edge vector (dense order, host) sampled at the centres of
   a gx x gy grid as edge2d.c:582-611 does: linear in x between x-edges
   (i, j), (i+1, j) and in y between y-edges (i, j), (i, j+1), with the
   reference's weights (1 - off, off) where off is the distance in units of
   length.  The neighbour past the last column (row) is the Bloch-periodic
   continuation xfactor * x(0, j) (yfactor * y(i, 0)); the reference reads a
   ghost slot there that none of its routines set. */
static void sample_field(const grid2d *gr, const field *v, int gx, int gy, field *vx, field *vy)
{
  const int nx = gr->nx, ny = gr->ny;
  const real hx = gr->hx, hy = gr->hy;
  const real hgx = (nx * hx) / gx, hgy = (ny * hy) / gy;
  const field *xe = v, *ye = v + (size_t)nx * ny;
  for (int jg = 0; jg < gy; jg++) {
    const int j = (jg + 0.5) * ny / gy;
    const double yoff = (jg + 0.5) * hgy - j * hy;
    for (int ig = 0; ig < gx; ig++) {
      const int i = (ig + 0.5) * nx / gx;
      const double xoff = (ig + 0.5) * hgx - i * hx;
      const field x1 = i + 1 < nx ? xe[(i + 1) + (size_t)j * nx] : gr->xfactor * xe[(size_t)j * nx];
      const field y1 = j + 1 < ny ? ye[i + (size_t)(j + 1) * nx] : gr->yfactor * ye[i];
      vx[ig * (size_t)gy + jg] = (1.0 - xoff) * xe[i + (size_t)j * nx] + xoff * x1;
      vy[ig * (size_t)gy + jg] = (1.0 - yoff) * ye[i + (size_t)j * nx] + yoff * y1;
    }
  }
}

void write_ncfile(ncfile *nf, int blochstep, const real *lambda, const edge2d **eigs)
{
#pragma omp single
  {
    if (blochstep < 0 || blochstep >= nf->blochsteps) nc_fail("Bloch step out of range");
    const grid2d *gr = eigs[0]->gr;
    const int gx = nf->graphx, gy = nf->graphy, ne = nf->eigenvalues;
    const double s[6] = {gr->xbloch, gr->ybloch, creal(gr->xfactor), cimag(gr->xfactor),
                         creal(gr->yfactor), cimag(gr->yfactor)};
    const int sid[6] = {nf->nc_xbloch, nf->nc_ybloch, nf->nc_xfactorr, nf->nc_xfactori, nf->nc_yfactorr,
                        nf->nc_yfactori};
    for (int k = 0; k < 6; k++) put_doubles(nf->file, nf->begin[sid[k]] + 8LL * blochstep, &s[k], 1);
    double *lam = (double *)malloc(sizeof(double) * ne);
    const size_t n = 2 * (size_t)gr->nx * gr->ny, g = (size_t)gx * gy;
    field *v = (field *)malloc(sizeof(field) * n), *vx = (field *)malloc(sizeof(field) * g),
          *vy = (field *)malloc(sizeof(field) * g);
    double *part = (double *)malloc(sizeof(double) * g);
    if (!lam || !v || !vx || !vy || !part) nc_fail("out of memory");
    for (int k = 0; k < ne; k++) lam[k] = lambda[k] - 1.0;
    put_doubles(nf->file, nf->begin[nf->nc_lambda] + 8LL * blochstep * ne, lam, ne);
    const int vid[4] = {nf->nc_vectorxr, nf->nc_vectorxi, nf->nc_vectoryr, nf->nc_vectoryi};
    for (int k = 0; k < ne; k++) {
      if (eigs[k]->gr != gr) nc_fail("eigenvectors on different grids");
      densefrom_edge2d(eigs[k], v);
      sample_field(gr, v, gx, gy, vx, vy);
      const long long off = 8LL * ((long long)blochstep * ne + k) * (long long)g;
      for (int c = 0; c < 4; c++) {
        const field *src = c < 2 ? vx : vy;
        for (size_t q = 0; q < g; q++) part[q] = (c & 1) ? cimag(src[q]) : creal(src[q]);
        put_doubles(nf->file, nf->begin[vid[c]] + off, part, g);
      }
    }
    free(lam);
    free(v);
    free(vx);
    free(vy);
    free(part);
  }
}

void cairodraw_edge2d(const edge2d *x, int gx, int gy, const epspattern *pat, const char *filename)
{
  (void)pat;
#pragma omp single
  {
    const grid2d *gr = x->gr;
    const size_t n = 2 * (size_t)gr->nx * gr->ny, g = (size_t)gx * gy;
    field *v = (field *)malloc(sizeof(field) * n), *vx = (field *)malloc(sizeof(field) * g),
          *vy = (field *)malloc(sizeof(field) * g);
    if (!v || !vx || !vy) nc_fail("out of memory");
    densefrom_edge2d(x, v);
    sample_field(gr, v, gx, gy, vx, vy);
    double mx = 0.0;
    for (size_t q = 0; q < g; q++) {
      const double a = sqrt(cabs(vx[q]) * cabs(vx[q]) + cabs(vy[q]) * cabs(vy[q]));
      if (a > mx) mx = a;
    }
    FILE *f = fopen(filename, "wb");
    if (f) {
      fprintf(f, "P5\n%d %d\n255\n", gx, gy);
      for (int jg = gy - 1; jg >= 0; jg--)
        for (int ig = 0; ig < gx; ig++) {
          const size_t q = ig * (size_t)gy + jg;
          const double a = sqrt(cabs(vx[q]) * cabs(vx[q]) + cabs(vy[q]) * cabs(vy[q]));
          fputc(mx > 0.0 ? (int)(255.0 * a / mx) : 0, f);
        }
      fclose(f);
    }
    free(v);
    free(vx);
    free(vy);
  }
}
