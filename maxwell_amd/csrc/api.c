/*
 * api.c -- the reference's module API (settings, basic, grid2d, linalg,
 * edge2d, node2d) over the gfx950 kernels of mx_kernels.hip.
 *
 * Contract kept from the reference:
 *  - vector functions are OpenMP collectives (edge2d.c uses `omp for` inside
 *    every routine): here one thread of the team launches the GPU work inside
 *    `omp single` and scalar results are broadcast with copyprivate, so the
 *    functions may be called by all threads of a team or serially;
 *  - dense functions (densematrix_*, denseto/densefrom, linalg) are called by
 *    one thread;
 *  - errors abort with a message, as the reference's asserts do.
 * Differences: coefficient arrays are device memory; the Gauss-Seidel sweeps
 * use four colours instead of the reference's row-parity line order
 * (edge2d.c:20 SMOOTHER_LINES), so smoother iterates differ while the
 * converged eigenvalues agree.
 */
#include <assert.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "basic.h"
#include "edge2d.h"
#include "grid2d.h"
#include "linalg.h"
#include "mx_dense.h"
#include "mxh.h"
#include "node2d.h"
#include "settings.h"

/* ================================================================ helpers */

static void die(const char *what, int err)
{
  fprintf(stderr, "maxwell_amd: %s failed: %s (%d)\n", what, err ? mxh_error_string(err) : "", err);
  abort();
}

#define HIPCK(call)                        \
  do {                                     \
    int _e = (call);                       \
    if (_e) die(#call, _e);                \
  } while (0)

static void ensure_init(void)
{
  static int done = 0;
  if (!done) {
    int e = mxh_init();
    if (e) die("mxh_init (is a GPU visible?)", e);
    done = 1;
  }
}

typedef struct {
  double *eta;   /* device copy of gr->eps (FP64) */
  double *work;  /* reduction scratch */
  double *out;   /* small device result buffer */
  field *tmp;    /* scratch edge vector */
  size_t ntmp;
} griddev;

static griddev *gdev(const grid2d *gr)
{
  griddev *g = (griddev *)gr->dev;
  if (!g) die("grid without device mirror", 0);
  return g;
}

static mxh_level level_of(const grid2d *gr)
{
  mxh_level l;
  l.nx = gr->nx;
  l.ny = gr->ny;
  l.hx = gr->hx;
  l.hy = gr->hy;
  return l;
}

static mxh_bloch bloch_of(const grid2d *gr)
{
  mxh_bloch b;
  memset(&b, 0, sizeof(b));
  b.xr = creal(gr->xfactor);
  b.xi = cimag(gr->xfactor);
  b.yr = creal(gr->yfactor);
  b.yi = cimag(gr->yfactor);
  b.vpp = 1;
  return b;
}

static mxh_vecs VV(const void *p, long long n)
{
  mxh_vecs v;
  v.p = (void *)p;
  v.stride = n;
  v.map = NULL;
  return v;
}

static mxh_coef coef(field a, field b)
{
  mxh_coef c;
  c.ar = creal(a);
  c.ai = cimag(a);
  c.br = creal(b);
  c.bi = cimag(b);
  c.per_vector = NULL;
  return c;
}

static long long en(const edge2d *x) { return 2LL * x->gr->nx * x->gr->ny; }
static long long nn(const node2d *x) { return (long long)x->gr->nx * x->gr->ny; }

static field *scratch_edge(const grid2d *gr)
{
  griddev *g = gdev(gr);
  size_t need = 2 * (size_t)gr->nx * gr->ny;
  if (g->ntmp < need) {
    mxh_free(g->tmp);
    g->tmp = (field *)mxh_malloc(sizeof(field) * need);
    if (!g->tmp) die("scratch allocation", 0);
    g->ntmp = need;
  }
  return g->tmp;
}

static field dot_dev(int prec, const grid2d *gr, long long n, const void *x, const void *y)
{
  griddev *g = gdev(gr);
  double r[2];
  HIPCK(mxh_dot(prec, n, 1, VV(x, n), VV(y, n), g->out, g->work));
  HIPCK(mxh_d2h(r, g->out, sizeof(r)));
  return r[0] + I * r[1];
}

static double norm_dev(const grid2d *gr, long long n, const void *x, int kind)
{
  griddev *g = gdev(gr);
  double r;
  HIPCK(mxh_norm(MXH_F64, n, 1, VV(x, n), kind, g->out, g->work));
  HIPCK(mxh_d2h(&r, g->out, sizeof(r)));
  return r;
}

/* ================================================================ settings */

real my_abs2(field x) { return creal(x) * creal(x) + cimag(x) * cimag(x); }

/* ================================================================ basic */

/* Deterministic team reductions (basic.c:11-161): every thread publishes its
   value, then every thread sums the slots in rank order. */
static void *team_slots(size_t bytes_per_thread)
{
  static void *slots;
#pragma omp single
  slots = calloc((size_t)omp_get_num_threads(), bytes_per_thread);
  return slots;
}

static void team_release(void *slots)
{
#pragma omp barrier
#pragma omp single
  free(slots);
}

real reduce_sum_real(real x)
{
  real *s = (real *)team_slots(sizeof(real));
  s[omp_get_thread_num()] = x;
#pragma omp barrier
  real r = 0.0;
  for (int k = 0; k < omp_get_num_threads(); k++) r += s[k];
  team_release(s);
  return r;
}

field reduce_sum_field(field x)
{
  field *s = (field *)team_slots(sizeof(field));
  s[omp_get_thread_num()] = x;
#pragma omp barrier
  field r = 0.0;
  for (int k = 0; k < omp_get_num_threads(); k++) r += s[k];
  team_release(s);
  return r;
}

real reduce_max_real(real x)
{
  real *s = (real *)team_slots(sizeof(real));
  s[omp_get_thread_num()] = x;
#pragma omp barrier
  real r = s[0];
  for (int k = 1; k < omp_get_num_threads(); k++)
    if (s[k] > r) r = s[k];
  team_release(s);
  return r;
}

field *bcast_fieldptr(field *a)
{
  static field *shared;
  if (a) shared = a;
#pragma omp barrier
  return shared;
}

/* ================================================================ grid2d */

static void upload_eps(grid2d *gr)
{
  griddev *g = gdev(gr);
  HIPCK(mxh_h2d(g->eta, gr->eps, sizeof(double) * (size_t)gr->nx * gr->ny));
}

grid2d *new_grid2d(int nx, int ny, real hx, real hy)
{
  ensure_init();
  grid2d *gr = (grid2d *)calloc(1, sizeof(grid2d));
  griddev *g = (griddev *)calloc(1, sizeof(griddev));
  assert(gr && g && nx > 0 && ny > 0);
  gr->hx = hx;
  gr->hy = hy;
  gr->nx = nx;
  gr->ny = ny;
  gr->xfactor = 1.0;
  gr->yfactor = 1.0;
  gr->eps = (real *)malloc(sizeof(real) * (size_t)nx * ny);
  assert(gr->eps);
  for (size_t k = 0; k < (size_t)nx * ny; k++) gr->eps[k] = 1.0;
  g->eta = (double *)mxh_malloc(sizeof(double) * (size_t)nx * ny);
  g->work = (double *)mxh_malloc(sizeof(double) * (513 * 4 + 64));
  g->out = (double *)mxh_malloc(sizeof(double) * 64);
  if (!g->eta || !g->work || !g->out) die("grid allocation", 0);
  gr->dev = g;
  upload_eps(gr);
  return gr;
}

void del_grid2d(grid2d *gr)
{
  griddev *g = (griddev *)gr->dev;
  if (g) {
    HIPCK(mxh_sync());
    mxh_free(g->eta);
    mxh_free(g->work);
    mxh_free(g->out);
    mxh_free(g->tmp);
    free(g);
  }
  free(gr->eps);
  gr->eps = NULL;
  free(gr);
}

void sync_grid2d(grid2d *gr) { upload_eps(gr); }

grid2d *coarsen_grid2d(const grid2d *fg)
{
  int nx = fg->nx / 2, ny = fg->ny / 2;
  assert(fg->nx == 2 * nx && fg->ny == 2 * ny);
  grid2d *gr = new_grid2d(nx, ny, 2.0 * fg->hx, 2.0 * fg->hy);
  gr->xfactor = fg->xfactor;
  gr->yfactor = fg->yfactor;
  gr->xbloch = fg->xbloch;
  gr->ybloch = fg->ybloch;
  const int fnx = fg->nx;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      gr->eps[i + j * nx] = 0.25 * (fg->eps[2 * i + 2 * j * fnx] + fg->eps[2 * i + 1 + 2 * j * fnx] +
                                    fg->eps[2 * i + (2 * j + 1) * fnx] + fg->eps[2 * i + 1 + (2 * j + 1) * fnx]);
  upload_eps(gr);
  return gr;
}

void draw2d_grid2d(const grid2d *gr, const char *filename)
{
#pragma omp single
  {
    FILE *f = fopen(filename, "wb");
    if (f) {
      double mx = gr->eps[0];
      for (size_t k = 0; k < (size_t)gr->nx * gr->ny; k++)
        if (gr->eps[k] > mx) mx = gr->eps[k];
      fprintf(f, "P5\n%d %d\n255\n", gr->nx, gr->ny);
      for (int j = gr->ny - 1; j >= 0; j--)
        for (int i = 0; i < gr->nx; i++) fputc((int)(255.0 * gr->eps[i + j * gr->nx] / mx), f);
      fclose(f);
    }
  }
}

epspattern *new_epspattern(real eps_base, int rectangles, int circles)
{
  epspattern *p = (epspattern *)malloc(sizeof(epspattern));
  assert(p);
  p->rectangle = (real *)malloc(sizeof(real) * 5 * (rectangles > 0 ? rectangles : 1));
  p->rectangles = rectangles;
  p->circle = (real *)malloc(sizeof(real) * 4 * (circles > 0 ? circles : 1));
  p->circles = circles;
  p->eps_base = eps_base;
  return p;
}

void del_epspattern(epspattern *pat)
{
  free(pat->circle);
  free(pat->rectangle);
  free(pat);
}

/* grid2d.c:171-218: 1/eps at the lower-left corner of each cell; rectangles
   are applied first, circles after. */
void setpattern_grid2d(const epspattern *pat, grid2d *gr)
{
  if (!pat) return;
  for (int j = 0; j < gr->ny; j++)
    for (int i = 0; i < gr->nx; i++) {
      double x = i * gr->hx, y = j * gr->hy, v = 1.0 / pat->eps_base;
      for (int k = 0; k < pat->rectangles; k++) {
        const real *r = pat->rectangle + 5 * k;
        if (r[1] <= x && x <= r[1] + r[3] && r[2] <= y && y <= r[2] + r[4]) v = 1.0 / r[0];
      }
      for (int k = 0; k < pat->circles; k++) {
        const real *c = pat->circle + 4 * k;
        if ((x - c[1]) * (x - c[1]) + (y - c[2]) * (y - c[2]) <= c[3] * c[3]) v = 1.0 / c[0];
      }
      gr->eps[i + j * gr->nx] = v;
    }
  upload_eps(gr);
}

/* ================================================================ linalg */

matrix *new_matrix(int rows, int cols)
{
  matrix *a = (matrix *)malloc(sizeof(matrix));
  assert(a);
  a->rows = rows;
  a->cols = cols;
  a->ld = rows;
  a->a = (field *)malloc(sizeof(field) * (size_t)(rows > 0 ? rows : 1) * (cols > 0 ? cols : 1));
  assert(a->a);
  return a;
}

void del_matrix(matrix *a)
{
  free(a->a);
  a->a = NULL;
  free(a);
}

static field *packed(const matrix *a)
{
  if (a->ld == a->rows) return a->a;
  field *p = (field *)malloc(sizeof(field) * (size_t)a->rows * a->cols);
  for (int j = 0; j < a->cols; j++)
    for (int i = 0; i < a->rows; i++) p[i + (size_t)j * a->rows] = a->a[i + (size_t)j * a->ld];
  return p;
}

static void unpack(matrix *a, field *p)
{
  if (p == a->a) return;
  for (int j = 0; j < a->cols; j++)
    for (int i = 0; i < a->rows; i++) a->a[i + (size_t)j * a->ld] = p[i + (size_t)j * a->rows];
  free(p);
}

int potrf_matrix(matrix *a)
{
  assert(a->rows == a->cols);
  field *p = packed(a);
  int info = mxd_potrf(a->rows, p);
  unpack(a, p);
  return info;
}

void potrsv_matrix(const matrix *a, field *b)
{
  assert(a->rows == a->cols);
  field *p = packed(a);
  mxd_potrs(a->rows, p, b);
  if (p != a->a) free(p);
}

int syev_matrix(matrix *a, real *lambda)
{
  assert(a->rows == a->cols);
  if (a->rows == 0) return 0;
  field *p = packed(a);
  int info = mxd_heev(a->rows, p, lambda);
  unpack(a, p);
  return info;
}

int sygv_matrix(matrix *a, matrix *b, real *lambda)
{
  assert(a->rows == a->cols && b->rows == b->cols && a->rows == b->rows);
  if (a->rows == 0) return 0;
  field *pa = packed(a), *pb = packed(b);
  int info = mxd_hegv(a->rows, pa, pb, lambda);
  unpack(a, pa);
  unpack(b, pb);
  return info;
}

/* ================================================================ edge2d */

edge2d *new_edge2d(const grid2d *gr)
{
  ensure_init();
  edge2d *x = (edge2d *)malloc(sizeof(edge2d));
  assert(x);
  size_t N = (size_t)gr->nx * gr->ny;
  x->gr = gr;
  x->x = (field *)mxh_malloc(sizeof(field) * 2 * N);
  if (!x->x) die("edge2d allocation", 0);
  x->y = x->x + N;
  return x;
}

void del_edge2d(edge2d *x)
{
  HIPCK(mxh_sync());
  mxh_free(x->x);
  x->x = x->y = NULL;
  x->gr = NULL;
  free(x);
}

size_t getdimension_edge2d(const edge2d *x) { return (size_t)2 * x->gr->nx * x->gr->ny; }

size_t getsize_edge2d(const edge2d *x)
{
  return sizeof(edge2d) + sizeof(field) * (size_t)2 * x->gr->nx * x->gr->ny;
}

void zero_edge2d(edge2d *x)
{
#pragma omp single
  HIPCK(mxh_memset(x->x, 0, sizeof(field) * en(x)));
}

void copy_edge2d(const edge2d *x, edge2d *y)
{
  assert(x->gr == y->gr);
#pragma omp single
  HIPCK(mxh_d2d(y->x, x->x, sizeof(field) * en(x)));
}

void swap_edge2d(edge2d *x, edge2d *y)
{
  assert(x->gr == y->gr);
#pragma omp single
  {
    field *a = x->x, *b = x->y;
    x->x = y->x;
    x->y = y->y;
    y->x = a;
    y->y = b;
  }
}

/* Uniform random entries in [-1,1] + i[-1,1] drawn with rand() in the
   reference's order (edge2d.c:161-191), then uploaded. */
void random_edge2d(edge2d *x)
{
#pragma omp single
  {
    long long n = en(x);
    field *h = (field *)malloc(sizeof(field) * n);
    for (long long k = 0; k < n; k++) {
      double re = 2.0 * rand() / RAND_MAX - 1.0;
      double im = 2.0 * rand() / RAND_MAX - 1.0;
      h[k] = re + I * im;
    }
    HIPCK(mxh_h2d(x->x, h, sizeof(field) * n));
    HIPCK(mxh_sync());
    free(h);
  }
}

void scale_edge2d(field alpha, edge2d *x)
{
#pragma omp single
  HIPCK(mxh_scale(MXH_F64, en(x), 1, VV(x->x, en(x)), creal(alpha), cimag(alpha)));
}

void add_edge2d(field alpha, const edge2d *x, edge2d *y)
{
  assert(x->gr == y->gr);
#pragma omp single
  HIPCK(mxh_axpy(MXH_F64, en(x), 1, VV(x->x, en(x)), VV(y->x, en(y)), creal(alpha), cimag(alpha)));
}

real norm2_edge2d(const edge2d *x)
{
  real r = 0.0;
#pragma omp single copyprivate(r)
  r = sqrt(norm_dev(x->gr, en(x), x->x, 0));
  return r;
}

real normmax_edge2d(const edge2d *x)
{
  real r = 0.0;
#pragma omp single copyprivate(r)
  r = norm_dev(x->gr, en(x), x->x, 1);
  return r;
}

field dotprod_edge2d(const edge2d *x, const edge2d *y)
{
  assert(x->gr == y->gr);
  field r = 0.0;
#pragma omp single copyprivate(r)
  r = dot_dev(MXH_F64, x->gr, en(x), x->x, y->x);
  return r;
}

/* edge2d.c:1240-1276: L2 products of the x and y components with the
   constant fields (Gamma point only): xprod = hy * sum of x-edge values,
   yprod = hx * sum of y-edge values. */
static void half_sums(const edge2d *x, field *sx, field *sy)
{
  const long long N = en(x) / 2;
  field *ones = scratch_edge(x->gr);
  field *h = (field *)malloc(sizeof(field) * 2 * N);
  if (!h) die("host allocation", 0);
  for (long long k = 0; k < 2 * N; k++) h[k] = (k < N) ? 1.0 : 0.0;
  HIPCK(mxh_h2d(ones, h, sizeof(field) * 2 * N));
  *sx = dot_dev(MXH_F64, x->gr, 2 * N, ones, x->x);
  for (long long k = 0; k < 2 * N; k++) h[k] = (k < N) ? 0.0 : 1.0;
  HIPCK(mxh_h2d(ones, h, sizeof(field) * 2 * N));
  *sy = dot_dev(MXH_F64, x->gr, 2 * N, ones, x->x);
  free(h);
}

void nullprod_edge2d(const edge2d *x, field *xprod, field *yprod)
{
  assert(x->gr->xfactor == 1.0 && x->gr->yfactor == 1.0);
  field a = 0.0, b = 0.0;
#pragma omp single copyprivate(a, b)
  {
    field sx, sy;
    half_sums(x, &sx, &sy);
    a = sx * x->gr->hy;
    b = sy * x->gr->hx;
  }
  *xprod = a;
  *yprod = b;
}

void nulladd_edge2d(field xalpha, field yalpha, edge2d *x)
{
  assert(x->gr->xfactor == 1.0 && x->gr->yfactor == 1.0);
#pragma omp single
  {
    const long long N = en(x) / 2;
    field *ones = scratch_edge(x->gr);
    field *h = (field *)malloc(sizeof(field) * 2 * N);
    for (long long k = 0; k < 2 * N; k++) h[k] = (k < N) ? xalpha * x->gr->hx : yalpha * x->gr->hy;
    HIPCK(mxh_h2d(ones, h, sizeof(field) * 2 * N));
    HIPCK(mxh_axpy(MXH_F64, 2 * N, 1, VV(ones, 2 * N), VV(x->x, 2 * N), 1.0, 0.0));
    HIPCK(mxh_sync());
    free(h);
  }
}

/* Subtract the mean of the x-edge and of the y-edge values (edge2d.c:1309-1333). */
void center_edge2d(edge2d *x)
{
  assert(x->gr->xfactor == 1.0 && x->gr->yfactor == 1.0);
#pragma omp single
  HIPCK(mxh_center(MXH_F64, en(x) / 2, 1, VV(x->x, en(x)), gdev(x->gr)->work));
}

void addeval_edge2d(field alpha, field beta, const edge2d *x, edge2d *y)
{
  assert(x->gr == y->gr);
#pragma omp single
  {
    mxh_level lv = level_of(x->gr);
    mxh_bloch bl = bloch_of(x->gr);
    mxh_coef cf = coef(alpha, beta);
    HIPCK(mxh_addeval(MXH_F64, &lv, &bl, 1, VV(x->x, en(x)), VV(y->x, en(y)), gdev(x->gr)->eta, &cf, 1));
  }
}

/* conj(y^H (alpha K + beta M) x), as the reference accumulates it
   (edge2d.c:1652-1959). */
field energyprod_edge2d(field alpha, field beta, const edge2d *x, const edge2d *y)
{
  assert(x->gr == y->gr);
  field r = 0.0;
#pragma omp single copyprivate(r)
  {
    mxh_level lv = level_of(x->gr);
    mxh_bloch bl = bloch_of(x->gr);
    mxh_coef cf = coef(alpha, beta);
    field *t = scratch_edge(x->gr);
    HIPCK(mxh_addeval(MXH_F64, &lv, &bl, 1, VV(x->x, en(x)), VV(t, en(x)), gdev(x->gr)->eta, &cf, 0));
    r = conj(dot_dev(MXH_F64, x->gr, en(x), y->x, t));
  }
  return r;
}

field massprod_edge2d(const edge2d *x, const edge2d *y)
{
  assert(x->gr == y->gr);
  field r = 0.0;
#pragma omp single copyprivate(r)
  {
    mxh_level lv = level_of(x->gr);
    mxh_bloch bl = bloch_of(x->gr);
    mxh_coef cf = coef(0.0, 1.0);
    field *t = scratch_edge(x->gr);
    HIPCK(mxh_addeval(MXH_F64, &lv, &bl, 1, VV(x->x, en(x)), VV(t, en(x)), gdev(x->gr)->eta, &cf, 0));
    r = conj(dot_dev(MXH_F64, x->gr, en(x), y->x, t));
  }
  return r;
}

static void gs_edge(field alpha, field beta, const edge2d *b, edge2d *x, int backward)
{
  mxh_level lv = level_of(x->gr);
  mxh_bloch bl = bloch_of(x->gr);
  for (int s = 0; s < 4; s++) {
    int c = backward ? 3 - s : s;
    HIPCK(mxh_afw_phase(MXH_F64, &lv, &bl, 1, VV(b->x, en(b)), VV(x->x, en(x)), gdev(x->gr)->eta, creal(alpha),
                        creal(beta), c));
  }
}

void gsforward_edge2d(field alpha, field beta, const edge2d *b, edge2d *x)
{
  assert(x->gr == b->gr);
#pragma omp single
  gs_edge(alpha, beta, b, x, 0);
}

void gsbackward_edge2d(field alpha, field beta, const edge2d *b, edge2d *x)
{
  assert(x->gr == b->gr);
#pragma omp single
  gs_edge(alpha, beta, b, x, 1);
}

void gssymm_edge2d(field alpha, field beta, const edge2d *b, edge2d *x)
{
  assert(x->gr == b->gr);
#pragma omp single
  {
    gs_edge(alpha, beta, b, x, 0);
    gs_edge(alpha, beta, b, x, 1);
  }
}

static void prolong_edge_impl(field alpha, const edge2d *c, edge2d *f)
{
  assert(f->gr->nx == 2 * c->gr->nx && f->gr->ny == 2 * c->gr->ny);
  mxh_level lv = level_of(f->gr);
  mxh_bloch bl = bloch_of(f->gr);
  if (cimag(alpha) == 0.0) {
    HIPCK(mxh_prolong(MXH_F64, &lv, &bl, 1, VV(c->x, en(c)), VV(f->x, en(f)), creal(alpha)));
  } else {
    field *t = scratch_edge(f->gr);
    HIPCK(mxh_memset(t, 0, sizeof(field) * en(f)));
    HIPCK(mxh_prolong(MXH_F64, &lv, &bl, 1, VV(c->x, en(c)), VV(t, en(f)), 1.0));
    HIPCK(mxh_axpy(MXH_F64, en(f), 1, VV(t, en(f)), VV(f->x, en(f)), creal(alpha), cimag(alpha)));
  }
}

void prolongation_edge2d(field alpha, const edge2d *c, edge2d *f)
{
#pragma omp single
  prolong_edge_impl(alpha, c, f);
}

void restriction_edge2d(field alpha, const edge2d *f, edge2d *c)
{
  assert(f->gr->nx == 2 * c->gr->nx && f->gr->ny == 2 * c->gr->ny);
#pragma omp single
  {
    mxh_level lv = level_of(f->gr);
    mxh_bloch bl = bloch_of(f->gr);
    if (cimag(alpha) == 0.0) {
      HIPCK(mxh_restrict(MXH_F64, &lv, &bl, 1, VV(f->x, en(f)), VV(c->x, en(c)), creal(alpha)));
    } else {
      field *t = scratch_edge(c->gr);
      HIPCK(mxh_memset(t, 0, sizeof(field) * en(c)));
      HIPCK(mxh_restrict(MXH_F64, &lv, &bl, 1, VV(f->x, en(f)), VV(t, en(c)), 1.0));
      HIPCK(mxh_axpy(MXH_F64, en(c), 1, VV(t, en(c)), VV(c->x, en(c)), creal(alpha), cimag(alpha)));
    }
  }
}

/* ---- coarse factor cache: the host factor is uploaded when it changes */
typedef struct {
  const matrix *key;
  uint64_t sum;
  int n;
  double *dev;
} factor_cache;

static uint64_t fnv(const void *p, size_t bytes)
{
  const unsigned char *c = (const unsigned char *)p;
  uint64_t h = 1469598103934665603ULL;
  for (size_t k = 0; k < bytes; k++) h = (h ^ c[k]) * 1099511628211ULL;
  return h;
}

static const double *device_factor(factor_cache *fc, const matrix *A)
{
  field *p = packed(A);
  size_t bytes = sizeof(field) * (size_t)A->rows * A->rows;
  uint64_t h = fnv(p, bytes);
  if (fc->key != A || fc->sum != h || fc->n != A->rows) {
    if (fc->n != A->rows) {
      mxh_free(fc->dev);
      fc->dev = (double *)mxh_malloc(bytes);
      if (!fc->dev) die("factor allocation", 0);
    }
    HIPCK(mxh_h2d(fc->dev, p, bytes));
    fc->key = A;
    fc->sum = h;
    fc->n = A->rows;
  }
  if (p != A->a) free(p);
  return fc->dev;
}

static factor_cache edge_fc, node_fc;

/* x = Ac^{-1} b on the coarsest level with the Cholesky factor in Ac
   (potrf_matrix output) */
static void edge_coarse_solve(const matrix *Ac, const edge2d *b, edge2d *x)
{
  assert(Ac->rows == 2 * x->gr->nx * x->gr->ny);
  if (Ac->rows <= 256) {
    const double *L0 = device_factor(&edge_fc, Ac);
    HIPCK(mxh_coarse_solve(MXH_F64, Ac->rows, L0, NULL, 1, NULL, 1, VV(b->x, en(b)), VV(x->x, en(x))));
  } else {
    field *h = (field *)malloc(sizeof(field) * Ac->rows);
    HIPCK(mxh_d2h(h, b->x, sizeof(field) * Ac->rows));
    potrsv_matrix(Ac, h);
    HIPCK(mxh_h2d(x->x, h, sizeof(field) * Ac->rows));
    HIPCK(mxh_sync());
    free(h);
  }
}

static void vcycle_edge_impl(int L, int nu, const matrix *Ac, edge2d **b, edge2d **x)
{
  for (int l = L; l > 0; l--) {
    for (int i = 0; i < nu; i++) gs_edge(1.0, 1.0, b[l], x[l], 0);
    mxh_level lv = level_of(x[l]->gr);
    mxh_bloch bl = bloch_of(x[l]->gr);
    HIPCK(mxh_resid_restrict(MXH_F64, &lv, &bl, 1, VV(b[l]->x, en(b[l])), VV(x[l]->x, en(x[l])),
                             gdev(x[l]->gr)->eta, VV(b[l - 1]->x, en(b[l - 1])), VV(x[l - 1]->x, en(x[l - 1]))));
  }
  if (Ac) {
    edge_coarse_solve(Ac, b[0], x[0]);
  } else {
    for (int i = 0; i < nu; i++) {
      gs_edge(1.0, 1.0, b[0], x[0], 0);
      gs_edge(1.0, 1.0, b[0], x[0], 1);
    }
  }
  for (int l = 1; l <= L; l++) {
    prolong_edge_impl(1.0, x[l - 1], x[l]);
    for (int i = 0; i < nu; i++) gs_edge(1.0, 1.0, b[l], x[l], 1);
  }
}

void vcycle_edge2d(int L, int nu, const matrix *Ac, edge2d **b, edge2d **x, edge2d **d)
{
  (void)d; /* residuals are formed inside the restriction kernel */
#pragma omp single
  vcycle_edge_impl(L, nu, Ac, b, x);
}

/* ---- pointwise edge Gauss-Seidel (edge2d.c:3892-5600).  The reference
   sweeps the edges row by row; here four colours (x-edges of even rows, of
   odd rows, y-edges of even columns, of odd columns) that are independent
   sets, so one colour is one launch. */
static void gs_simple(field alpha, field beta, const edge2d *b, edge2d *x, int backward)
{
  mxh_level lv = level_of(x->gr);
  mxh_bloch bl = bloch_of(x->gr);
  mxh_coef cf = coef(alpha, beta);
  for (int s = 0; s < 4; s++)
    HIPCK(mxh_edge_gs_phase(MXH_F64, &lv, &bl, 1, VV(b->x, en(b)), VV(x->x, en(x)), gdev(x->gr)->eta, &cf,
                            backward ? 3 - s : s));
}

void gsforward_simple_edge2d(field alpha, field beta, const edge2d *b, edge2d *x)
{
  assert(x->gr == b->gr);
#pragma omp single
  gs_simple(alpha, beta, b, x, 0);
}

void gsbackward_simple_edge2d(field alpha, field beta, const edge2d *b, edge2d *x)
{
  assert(x->gr == b->gr);
#pragma omp single
  gs_simple(alpha, beta, b, x, 1);
}

void gssymm_simple_edge2d(field alpha, field beta, const edge2d *b, edge2d *x)
{
  assert(x->gr == b->gr);
#pragma omp single
  {
    gs_simple(alpha, beta, b, x, 0);
    gs_simple(alpha, beta, b, x, 1);
  }
}

/* ---- Hiptmair's V-cycle for K + M (edge2d.c:5757-5823): on every level the
   pointwise edge smoother, then the same smoother for the gradient part,
   i.e. nodal Gauss-Seidel for G^H (K + M) G = G^H M G (the Q1 Laplacian) on
   the residual mapped by G^H, corrected back through G.  Coarsest level:
   Cholesky factor Ac or nu symmetric sweeps.  Built from the collective
   vector functions, so it is itself a collective of the calling team. */
void hcycle_edge2d(int L, int nu, const matrix *Ac, edge2d **b, edge2d **x, edge2d **d, node2d **bg,
                   node2d **xg)
{
  for (int l = L; l > 0; l--) {
    for (int i = 0; i < nu; i++) gsforward_simple_edge2d(1.0, 1.0, b[l], x[l]);
    copy_edge2d(b[l], d[l]);
    addeval_edge2d(-1.0, -1.0, x[l], d[l]);
    zero_node2d(bg[l]);
    adjgradient_node2d(1.0, d[l], bg[l]);
    zero_node2d(xg[l]);
    for (int i = 0; i < nu; i++) gsforward_node2d(1.0, bg[l], xg[l]);
    gradient_node2d(1.0, xg[l], x[l]);
    copy_edge2d(b[l], d[l]);
    addeval_edge2d(-1.0, -1.0, x[l], d[l]);
    zero_edge2d(b[l - 1]);
    restriction_edge2d(1.0, d[l], b[l - 1]);
    zero_edge2d(x[l - 1]);
  }
  if (Ac) {
#pragma omp single
    edge_coarse_solve(Ac, b[0], x[0]);
  } else {
    for (int i = 0; i < nu; i++) gssymm_simple_edge2d(1.0, 1.0, b[0], x[0]);
  }
  for (int l = 1; l <= L; l++) {
    prolongation_edge2d(1.0, x[l - 1], x[l]);
    copy_edge2d(b[l], d[l]);
    addeval_edge2d(-1.0, -1.0, x[l], d[l]);
    zero_node2d(bg[l]);
    adjgradient_node2d(1.0, d[l], bg[l]);
    zero_node2d(xg[l]);
    for (int i = 0; i < nu; i++) gsbackward_node2d(1.0, bg[l], xg[l]);
    gradient_node2d(1.0, xg[l], x[l]);
    for (int i = 0; i < nu; i++) gsbackward_simple_edge2d(1.0, 1.0, b[l], x[l]);
  }
}

/* ---- conjugate gradients for (K + M) x = b preconditioned by one AFW
   V(nu, nu) cycle without a coarse matrix (edge2d.c:5826-5868): init sets
   r = b - A x, q = V r, p = q; a step updates x, r (recursively, as the
   reference does by default) and p.  r, q, d: level vectors of the cycle. */
void mgcginit_edge2d(int L, int nu, const edge2d *b, edge2d *x, edge2d *p, edge2d *a, edge2d **r, edge2d **q,
                     edge2d **d)
{
  (void)a;
  copy_edge2d(b, r[L]);
  addeval_edge2d(-1.0, -1.0, x, r[L]);
  zero_edge2d(q[L]);
  vcycle_edge2d(L, nu, NULL, r, q, d);
  copy_edge2d(q[L], p);
}

void mgcgstep_edge2d(int L, int nu, const edge2d *b, edge2d *x, edge2d *p, edge2d *a, edge2d **r, edge2d **q,
                     edge2d **d)
{
  (void)b;
  zero_edge2d(a);
  addeval_edge2d(1.0, 1.0, p, a);
  const field gamma = dotprod_edge2d(p, a);
  const field lambda = dotprod_edge2d(p, r[L]) / gamma;
  add_edge2d(lambda, p, x);
  add_edge2d(-lambda, a, r[L]);
  zero_edge2d(q[L]);
  vcycle_edge2d(L, nu, NULL, r, q, d);
  const field mu = dotprod_edge2d(a, q[L]) / gamma;
  scale_edge2d(-mu, p);
  add_edge2d(1.0, q[L], p);
}

void densematrix_edge2d(field alpha, field beta, const grid2d *gr, matrix *a)
{
  assert(a->rows == a->cols && a->rows == 2 * gr->nx * gr->ny);
  mxd_dense_edge(gr->nx, gr->ny, gr->hx, gr->hy, gr->xfactor, gr->yfactor, gr->eps, alpha, beta, a->a, a->ld);
}

void denseto_edge2d(const field *x, edge2d *y)
{
  HIPCK(mxh_h2d(y->x, x, sizeof(field) * en(y)));
  HIPCK(mxh_sync());
}

void densefrom_edge2d(const edge2d *x, field *y) { HIPCK(mxh_d2h(y, x->x, sizeof(field) * en(x))); }

/* ---- Householder M-orthonormalisation (edge2d.c:6154-6296).  The unit
   vectors sit on x-edges of odd rows: idx -> (i, j) with j = 1, 3, ... */
void unitindices_edge2d(int idx, const edge2d *x, int *i, int *j)
{
  int nx = x->gr->nx, ny = x->gr->ny;
  int ii = idx, jj;
  for (jj = 1; jj < ny && ii >= nx; jj += 2) ii -= nx;
  assert(jj < ny);
  *i = ii;
  *j = jj;
}

static double unit_scale(const grid2d *gr) { return sqrt(0.25 / (gr->hy / gr->hx / 6.0)); }

void unit_edge2d(int idx, edge2d *x)
{
  int i, j;
  unitindices_edge2d(idx, x, &i, &j);
#pragma omp single
  {
    field v = unit_scale(x->gr);
    HIPCK(mxh_memset(x->x, 0, sizeof(field) * en(x)));
    HIPCK(mxh_h2d(x->x + i + (size_t)j * x->gr->nx, &v, sizeof(field)));
    HIPCK(mxh_sync());
  }
}

field massprod_xunit_edge2d(int i, int j, const edge2d *y)
{
  const int nx = y->gr->nx, ny = y->gr->ny;
  assert(0 <= i && i < nx && 0 < j && j < ny);
  field r = 0.0;
#pragma omp single copyprivate(r)
  {
    double mx = y->gr->hy / y->gr->hx / 6.0;
    field v[3];
    HIPCK(mxh_d2h(&v[0], y->x + i + (size_t)(j - 1) * nx, sizeof(field)));
    HIPCK(mxh_d2h(&v[1], y->x + i + (size_t)j * nx, sizeof(field)));
    if (j + 1 < ny) HIPCK(mxh_d2h(&v[2], y->x + i + (size_t)(j + 1) * nx, sizeof(field)));
    else {
      HIPCK(mxh_d2h(&v[2], y->x + i, sizeof(field)));
      v[2] *= y->gr->yfactor;
    }
    r = sqrt(0.25 * mx) * (4.0 * v[1] + v[0] + v[2]);
  }
  return r;
}

void add_xunit_edge2d(field alpha, int i, int j, edge2d *y)
{
  assert(0 <= i && i < y->gr->nx && 0 < j && j < y->gr->ny);
#pragma omp single
  {
    field v;
    field *p = y->x + i + (size_t)j * y->gr->nx;
    HIPCK(mxh_d2h(&v, p, sizeof(field)));
    v += alpha * unit_scale(y->gr);
    HIPCK(mxh_h2d(p, &v, sizeof(field)));
    HIPCK(mxh_sync());
  }
}

void buildhouseholder_edge2d(int idx, edge2d *x, field *tau)
{
  int i, j;
  unitindices_edge2d(idx, x, &i, &j);
  real alpha = sqrt(creal(massprod_edge2d(x, x)));
  if (alpha == 0.0) {
    *tau = 0.0;
    return;
  }
  field beta = massprod_xunit_edge2d(i, j, x);
  if (creal(beta) < 0.0) alpha = -alpha;
  add_xunit_edge2d(alpha, i, j, x);
  *tau = 1.0 / (alpha * alpha + beta * alpha);
}

void applyhouseholder_edge2d(const edge2d *v, field tau, edge2d *x)
{
  field beta = massprod_edge2d(v, x);
  add_edge2d(-tau * beta, v, x);
}

/* The whole orthonormalisation runs on the device as one sequence of
   launches over a contiguous copy of the k vectors (mxh_orthonormalize: per
   Householder step a batched M-product row, one rank-one update of all later
   vectors and their unit-entry fix; the outputs H_0 ... H_i unit_i are built
   for all i together), with no host round trip until tau is copied back.
   Outputs go to x[0..k-1]; v (scratch in the reference) receives x[0]'s
   output, as the reference's last back-substitution leaves it. */
void orthonormalize_edge2d(int k, edge2d **x, field *tau, edge2d *v)
{
  assert(k >= 1);
#pragma omp single
  {
    const grid2d *gr = x[0]->gr;
    const size_t n = en(x[0]);
    mxh_level lv = level_of(gr);
    field *blk = (field *)mxh_malloc(sizeof(field) * n * (2 * (size_t)k + 1));
    double *dtau = (double *)mxh_malloc(sizeof(double) * 2 * (size_t)k);
    double *work = (double *)mxh_malloc(mxh_orthonormalize_work_bytes(k));
    if (!blk || !dtau || !work) die("orthonormalize allocation", 0);
    field *X = blk, *Vv = blk + (size_t)k * n, *tmp = blk + 2 * (size_t)k * n;
    for (int i = 0; i < k; i++) {
      assert(x[i]->gr == gr);
      HIPCK(mxh_d2d(X + (size_t)i * n, x[i]->x, sizeof(field) * n));
    }
    HIPCK(mxh_orthonormalize(&lv, creal(gr->xfactor), cimag(gr->xfactor), creal(gr->yfactor), cimag(gr->yfactor),
                             gdev(gr)->eta, k, (double *)X, (double *)Vv, (double *)tmp, dtau, work));
    for (int i = 0; i < k; i++) HIPCK(mxh_d2d(x[i]->x, Vv + (size_t)i * n, sizeof(field) * n));
    if (v) HIPCK(mxh_d2d(v->x, Vv, sizeof(field) * n));
    HIPCK(mxh_d2h(tau, dtau, sizeof(field) * (size_t)k));
    mxh_free(blk);
    mxh_free(dtau);
    mxh_free(work);
  }
}

/* ================================================================ node2d */

node2d *new_node2d(const grid2d *gr)
{
  ensure_init();
  node2d *x = (node2d *)malloc(sizeof(node2d));
  assert(x);
  x->gr = gr;
  x->v = (field *)mxh_malloc(sizeof(field) * (size_t)gr->nx * gr->ny);
  if (!x->v) die("node2d allocation", 0);
  return x;
}

void del_node2d(node2d *x)
{
  HIPCK(mxh_sync());
  mxh_free(x->v);
  x->v = NULL;
  x->gr = NULL;
  free(x);
}

void zero_node2d(node2d *x)
{
#pragma omp single
  HIPCK(mxh_memset(x->v, 0, sizeof(field) * nn(x)));
}

void copy_node2d(const node2d *x, node2d *y)
{
  assert(x->gr == y->gr);
#pragma omp single
  HIPCK(mxh_d2d(y->v, x->v, sizeof(field) * nn(x)));
}

void random_node2d(node2d *x)
{
#pragma omp single
  {
    long long n = nn(x);
    field *h = (field *)malloc(sizeof(field) * n);
    for (long long k = 0; k < n; k++) {
      double re = 2.0 * rand() / RAND_MAX - 1.0;
      double im = 2.0 * rand() / RAND_MAX - 1.0;
      h[k] = re + I * im;
    }
    HIPCK(mxh_h2d(x->v, h, sizeof(field) * n));
    HIPCK(mxh_sync());
    free(h);
  }
}

void scale_node2d(field alpha, node2d *x)
{
#pragma omp single
  HIPCK(mxh_scale(MXH_F64, nn(x), 1, VV(x->v, nn(x)), creal(alpha), cimag(alpha)));
}

void add_node2d(field alpha, const node2d *x, node2d *y)
{
  assert(x->gr == y->gr);
#pragma omp single
  HIPCK(mxh_axpy(MXH_F64, nn(x), 1, VV(x->v, nn(x)), VV(y->v, nn(y)), creal(alpha), cimag(alpha)));
}

real norm2_node2d(const node2d *x)
{
  real r = 0.0;
#pragma omp single copyprivate(r)
  r = sqrt(norm_dev(x->gr, nn(x), x->v, 0));
  return r;
}

real normmax_node2d(const node2d *x)
{
  real r = 0.0;
#pragma omp single copyprivate(r)
  r = norm_dev(x->gr, nn(x), x->v, 1);
  return r;
}

field dotprod_node2d(const node2d *x, const node2d *y)
{
  assert(x->gr == y->gr);
  field r = 0.0;
#pragma omp single copyprivate(r)
  r = dot_dev(MXH_F64, x->gr, nn(x), x->v, y->v);
  return r;
}

/* node2d.c:490-514: subtract the mean */
void center_node2d(node2d *x)
{
#pragma omp single
  {
    long long n = nn(x);
    field *ones = scratch_edge(x->gr);
    field *h = (field *)malloc(sizeof(field) * n);
    for (long long k = 0; k < n; k++) h[k] = 1.0;
    HIPCK(mxh_h2d(ones, h, sizeof(field) * n));
    field s = dot_dev(MXH_F64, x->gr, n, ones, x->v) / (double)n;
    HIPCK(mxh_axpy(MXH_F64, n, 1, VV(ones, n), VV(x->v, n), -creal(s), -cimag(s)));
    HIPCK(mxh_sync());
    free(h);
  }
}

void addeval_node2d(field alpha, const node2d *x, node2d *y)
{
  assert(x->gr == y->gr);
#pragma omp single
  {
    mxh_level lv = level_of(x->gr);
    mxh_bloch bl = bloch_of(x->gr);
    HIPCK(mxh_node_addeval(MXH_F64, &lv, &bl, 1, VV(x->v, nn(x)), VV(y->v, nn(y)), creal(alpha), cimag(alpha), 1));
  }
}

static void gs_node(field alpha, const node2d *b, node2d *x, int backward)
{
  mxh_level lv = level_of(x->gr);
  mxh_bloch bl = bloch_of(x->gr);
  for (int s = 0; s < 4; s++) {
    int c = backward ? 3 - s : s;
    HIPCK(mxh_node_gs_phase(MXH_F64, &lv, &bl, 1, VV(b->v, nn(b)), VV(x->v, nn(x)), creal(alpha), c));
  }
}

void gsforward_node2d(field alpha, const node2d *b, node2d *x)
{
#pragma omp single
  gs_node(alpha, b, x, 0);
}

void gsbackward_node2d(field alpha, const node2d *b, node2d *x)
{
#pragma omp single
  gs_node(alpha, b, x, 1);
}

void gssymm_node2d(field alpha, const node2d *b, node2d *x)
{
#pragma omp single
  {
    gs_node(alpha, b, x, 0);
    gs_node(alpha, b, x, 1);
  }
}

void prolongation_node2d(field alpha, const node2d *c, node2d *f)
{
  assert(f->gr->nx == 2 * c->gr->nx && f->gr->ny == 2 * c->gr->ny);
#pragma omp single
  {
    mxh_level lv = level_of(f->gr);
    mxh_bloch bl = bloch_of(f->gr);
    HIPCK(mxh_node_prolong(MXH_F64, &lv, &bl, 1, VV(c->v, nn(c)), VV(f->v, nn(f)), creal(alpha), cimag(alpha)));
  }
}

void restriction_node2d(field alpha, const node2d *f, node2d *c)
{
  assert(f->gr->nx == 2 * c->gr->nx && f->gr->ny == 2 * c->gr->ny);
#pragma omp single
  {
    mxh_level lv = level_of(f->gr);
    mxh_bloch bl = bloch_of(f->gr);
    HIPCK(mxh_node_restrict(MXH_F64, &lv, &bl, 1, VV(f->v, nn(f)), VV(c->v, nn(c)), creal(alpha), cimag(alpha)));
  }
}

void gradient_node2d(field alpha, const node2d *x, edge2d *y)
{
  assert(y->gr == x->gr);
#pragma omp single
  {
    mxh_level lv = level_of(x->gr);
    mxh_bloch bl = bloch_of(x->gr);
    HIPCK(mxh_gradient(MXH_F64, &lv, &bl, 1, VV(x->v, nn(x)), VV(y->x, en(y)), creal(alpha), cimag(alpha)));
  }
}

void adjgradient_node2d(field alpha, const edge2d *x, node2d *y)
{
  assert(y->gr == x->gr);
#pragma omp single
  {
    mxh_level lv = level_of(x->gr);
    mxh_bloch bl = bloch_of(x->gr);
    HIPCK(mxh_adjgradient(MXH_F64, &lv, &bl, 1, VV(x->x, en(x)), VV(y->v, nn(y)), creal(alpha), cimag(alpha)));
  }
}

static void vcycle_node_impl(int L, int nu, const matrix *Ac, node2d **b, node2d **x)
{
  for (int l = L; l > 0; l--) {
    for (int i = 0; i < nu; i++) gs_node(1.0, b[l], x[l], 0);
    mxh_level lv = level_of(x[l]->gr);
    mxh_bloch bl = bloch_of(x[l]->gr);
    HIPCK(mxh_node_resid_restrict(MXH_F64, &lv, &bl, 1, VV(b[l]->v, nn(b[l])), VV(x[l]->v, nn(x[l])),
                                  VV(b[l - 1]->v, nn(b[l - 1])), VV(x[l - 1]->v, nn(x[l - 1]))));
  }
  if (Ac) {
    assert(Ac->rows == x[0]->gr->nx * x[0]->gr->ny);
    if (Ac->rows <= 256) {
      const double *L0 = device_factor(&node_fc, Ac);
      HIPCK(mxh_coarse_solve(MXH_F64, Ac->rows, L0, NULL, 1, NULL, 1, VV(b[0]->v, nn(b[0])), VV(x[0]->v, nn(x[0]))));
    } else {
      field *h = (field *)malloc(sizeof(field) * Ac->rows);
      HIPCK(mxh_d2h(h, b[0]->v, sizeof(field) * Ac->rows));
      potrsv_matrix(Ac, h);
      HIPCK(mxh_h2d(x[0]->v, h, sizeof(field) * Ac->rows));
      HIPCK(mxh_sync());
      free(h);
    }
  } else {
    for (int i = 0; i < nu; i++) {
      gs_node(1.0, b[0], x[0], 0);
      gs_node(1.0, b[0], x[0], 1);
    }
  }
  for (int l = 1; l <= L; l++) {
    mxh_level lv = level_of(x[l]->gr);
    mxh_bloch bl = bloch_of(x[l]->gr);
    HIPCK(mxh_node_prolong(MXH_F64, &lv, &bl, 1, VV(x[l - 1]->v, nn(x[l - 1])), VV(x[l]->v, nn(x[l])), 1.0, 0.0));
    for (int i = 0; i < nu; i++) gs_node(1.0, b[l], x[l], 1);
  }
}

void vcycle_node2d(int L, int nu, const matrix *Ac, node2d **b, node2d **x, node2d **d)
{
  (void)d;
#pragma omp single
  vcycle_node_impl(L, nu, Ac, b, x);
}

/* ---- conjugate gradients for the Q1 Laplacian preconditioned by one nodal
   V(nu, nu) cycle without a coarse matrix (node2d.c:1392-1427). */
void mgcginit_node2d(int L, int nu, const node2d *b, node2d *x, node2d *p, node2d *a, node2d **r, node2d **q,
                     node2d **d)
{
  (void)a;
  copy_node2d(b, r[L]);
  addeval_node2d(-1.0, x, r[L]);
  zero_node2d(q[L]);
  vcycle_node2d(L, nu, NULL, r, q, d);
  copy_node2d(q[L], p);
}

void mgcgstep_node2d(int L, int nu, const node2d *b, node2d *x, node2d *p, node2d *a, node2d **r, node2d **q,
                     node2d **d)
{
  (void)b;
  zero_node2d(a);
  addeval_node2d(1.0, p, a);
  const field gamma = dotprod_node2d(p, a);
  const field lambda = dotprod_node2d(p, r[L]) / gamma;
  add_node2d(lambda, p, x);
  add_node2d(-lambda, a, r[L]);
  zero_node2d(q[L]);
  vcycle_node2d(L, nu, NULL, r, q, d);
  const field mu = dotprod_node2d(a, q[L]) / gamma;
  scale_node2d(-mu, p);
  add_node2d(1.0, q[L], p);
}

void densematrix_node2d(bool fixnull, const grid2d *gr, matrix *a)
{
  assert(a->rows == a->cols && a->rows == gr->nx * gr->ny);
  mxd_dense_node(gr->nx, gr->ny, gr->hx, gr->hy, gr->xfactor, gr->yfactor, fixnull, a->a, a->ld);
}

void denseto_node2d(const field *x, node2d *y)
{
  HIPCK(mxh_h2d(y->v, x, sizeof(field) * nn(y)));
  HIPCK(mxh_sync());
}

void densefrom_node2d(const node2d *x, field *y) { HIPCK(mxh_d2h(y, x->v, sizeof(field) * nn(x))); }

/* ================================================================ PINVIT */

/* edge2d.c:6298-6328: b = (K + (1-lambda) M) e; x = V^prec(b); e -= x;
   e -= G V_node^grad(G^H M e). */
void pinvit_edge2d(int l, int smoother_steps, int prec_steps, int gradient_steps, real lambda, edge2d *e,
                   const matrix *Ae, edge2d **b, edge2d **x, edge2d **d, const matrix *An, node2d **bg,
                   node2d **xg, node2d **dg)
{
  (void)d;
  (void)dg;
#pragma omp single
  {
    const grid2d *gr = e->gr;
    mxh_level lv = level_of(gr);
    mxh_bloch bl = bloch_of(gr);
    mxh_coef cf = coef(1.0, 1.0 - lambda);
    const long long n = en(e), N = n / 2;
    HIPCK(mxh_addeval(MXH_F64, &lv, &bl, 1, VV(e->x, n), VV(b[l]->x, n), gdev(gr)->eta, &cf, 0));
    HIPCK(mxh_memset(x[l]->x, 0, sizeof(field) * n));
    for (int j = 0; j < prec_steps; j++) vcycle_edge_impl(l, smoother_steps, Ae, b, x);
    /* lifting right-hand side of u = e - x, then e <- u - G xg in place (u
       is formed from e and x by both kernels, never stored) */
    HIPCK(mxh_lift_rhs(MXH_F64, &lv, &bl, 1, VV(e->x, n), VV(x[l]->x, n), VV(NULL, n), VV(bg[l]->v, N),
                       VV(xg[l]->v, N), NULL, NULL));
    for (int j = 0; j < gradient_steps; j++) vcycle_node_impl(l, smoother_steps, An, bg, xg);
    HIPCK(mxh_lift_apply(MXH_F64, &lv, &bl, 1, VV(xg[l]->v, N), VV(e->x, n), VV(x[l]->x, n), VV(e->x, n), -1.0,
                         NULL, NULL));
  }
}
