/*
 * maxwell_oracle.h -- CPU restatement (complex FP64) of the reference's Bloch
 * eigensolver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test suite, the smoke check and the CPU
 * baseline leg of the benchmark may load this library, and only to check
 * results or as the timed CPU baseline.  The product library never links it.
 *
 * Layout (identical to the reference's dense ordering, edge2d.c:5870-5905 and
 * edge2d.c:6112-6152): an edge vector of an nx x ny periodic grid holds
 * 2*nx*ny complex values, x-edge (i,j) at [i + j*nx] and y-edge (i,j) at
 * [nx*ny + i + j*nx].  x-edge (i,j) runs from node (i,j) to node (i+1,j),
 * y-edge (i,j) from node (i,j) to node (i,j+1).  Cell (i,j) has bottom x-edge
 * (i,j), top x-edge (i,j+1), left y-edge (i,j), right y-edge (i+1,j).  A node
 * vector holds nx*ny values, node (i,j) at [i + j*nx].  eta[i + j*nx] is
 * 1/eps of cell (i,j) (grid2d.h:33).
 *
 * Complex arrays are passed as double* with interleaved (re, im) pairs, which
 * is the memory layout of C99 `double complex`.
 */
#ifndef MAXWELL_ORACLE_H
#define MAXWELL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny;
  double hx, hy;
  double xf[2];          /* Bloch factor e^{i kx} (re, im) */
  double yf[2];          /* Bloch factor e^{i ky} (re, im) */
  const double *eta;     /* 1/eps per cell */
} og_grid;

/* ---- grid2d (grid2d.c:42-68, 171-218) ---- */
void og_setpattern(int nx, int ny, double hx, double hy, double eps_base,
                   int nrect, const double *rect, int ncirc, const double *circ,
                   double *eta);
void og_coarsen(int nx, int ny, const double *eta_f, double *eta_c);

/* ---- edge2d operators ---- */
void og_addeval(const og_grid *g, double ar, double ai, double br, double bi,
                const double *x, double *y, int quirk);
void og_energyprod(const og_grid *g, double ar, double ai, double br, double bi,
                   const double *x, const double *y, int quirk, double *out);
void og_massprod(const og_grid *g, const double *x, const double *y, double *out);
void og_dotprod(int n, const double *x, const double *y, double *out);
double og_norm2(int n, const double *x);
void og_center(const og_grid *g, double *x);

/* ---- AFW vertex-patch smoother ---- */
void og_patch_solve(const og_grid *g, double alpha, double beta,
                    const double *b, double *x, int i, int j);
void og_gs_edge(const og_grid *g, double alpha, double beta,
                const double *b, double *x, int backward, int colour_order);

/* ---- transfers ---- */
void og_prolong_edge(const og_grid *gf, double ar, double ai,
                     const double *c, double *f);
void og_restrict_edge(const og_grid *gf, double ar, double ai,
                      const double *f, double *c);

/* ---- dense ---- */
void og_dense_edge(const og_grid *g, double ar, double ai, double br, double bi,
                   double *a);
void og_dense_node(const og_grid *g, int fixnull, double *a);
int og_potrf(int n, double *a);
void og_potrs(int n, const double *l, double *b);
int og_hegv(int n, double *a, double *b, double *w);

/* ---- node2d ---- */
void og_addeval_node(const og_grid *g, double ar, double ai,
                     const double *x, double *y);
void og_gs_node(const og_grid *g, double alpha, const double *b, double *x,
                int backward, int colour_order);
void og_prolong_node(const og_grid *gf, double ar, double ai,
                     const double *c, double *f);
void og_restrict_node(const og_grid *gf, double ar, double ai,
                      const double *f, double *c);
void og_gradient(const og_grid *g, double ar, double ai,
                 const double *v, double *e);
void og_adjgradient(const og_grid *g, double ar, double ai,
                    const double *e, double *v);

/* ---- multigrid hierarchy and eigensolver ---- */
typedef struct og_mg og_mg;

og_mg *og_mg_new(int n0, int L, const double *eta_fine, int colour_order,
                 int quirk);
void og_mg_free(og_mg *mg);
void og_mg_set_bloch(og_mg *mg, double kx, double ky);
void og_mg_vcycle_edge(og_mg *mg, int nu, const double *b, double *x);
void og_mg_vcycle_node(og_mg *mg, int nu, const double *b, double *x);
const double *og_mg_eta(const og_mg *mg, int l);

/* One PINVIT vector step (edge2d.c:6298-6328 with the Gamma-point centring of
   test_block2d.c:217-220).  `lam` is the current Ritz value of K+M. */
void og_pinvit_vector(og_mg *mg, int prec_steps, int gradient_steps,
                      double lam, double *e);

/* Gradient lifting alone: e -= G * MG_node^{steps}(G^H M e). */
void og_lift(og_mg *mg, int gradient_steps, double *e);

/* Block PINVIT + Ritz at the Bloch parameter set by og_mg_set_bloch.
 * E: blk vectors (in: start block, out: Ritz vectors).  Stops when the
 * relative residual |K e - lam M e| / (|lam| |M e|) of the first neig Ritz
 * pairs is <= tol, or after maxit block steps.  lam_out: blk Ritz values of
 * K (not K+M), relres_out/absres_out: blk residuals.  Returns the number of
 * block steps taken. */
int og_block_solve(og_mg *mg, int neig, int blk, double tol, int maxit,
                   int prec_steps, int gradient_steps, double *E,
                   double *lam_out, double *relres_out, double *absres_out);

/* Adjoint: d lambda / d eta_c = |curl e|_c^2 / (hx hy) / (e^H M e) for one
   eigenvector, on the finest grid of `g`. */
void og_eta_gradient(const og_grid *g, const double *e, double *grad);

#ifdef __cplusplus
}
#endif

#endif
