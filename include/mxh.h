/*
 * mxh.h -- thin C-ABI between the C host layer and the gfx950 HIP kernels.
 *
 * Only plain data crosses this boundary: ints, doubles, device pointers and
 * small POD structs.  Every launcher enqueues on the library's stream and
 * returns 0 or a HIP error code; none of them synchronises unless its name says
 * so.
 *
 * Storage (device memory): an edge vector of an nx x ny periodic grid is
 * 2*nx*ny complex values, x-edges (i,j) at [i + j*nx] then y-edges at
 * [nx*ny + i + j*nx]; complex values are (re, im) pairs of the precision
 * `prec` (MXH_F64: double, MXH_F32: float).  A block of nv vectors is nv such
 * vectors `stride` complex values apart.  Node vectors hold nx*ny values.
 * eta (1/eps per cell) is stored in the vector precision.
 */
#ifndef MXH_H
#define MXH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MXH_F64 = 0, MXH_F32 = 1 };

typedef struct {
  int nx, ny;
  double hx, hy;
} mxh_level;

/* Bloch factors xf = e^{i kx}, yf = e^{i ky}.  If per_problem is non-null it
   points to device doubles (xf.re, xf.im, yf.re, yf.im) per problem and vector
   v belongs to problem g / vpp with g = vmap ? vmap[v] : v (vmap: device ints,
   the batch-wide index of a compacted vector); otherwise the inline factors
   hold for all. */
typedef struct {
  double xr, xi, yr, yi;
  const double *per_problem;
  int vpp;
  const int *vmap;
} mxh_bloch;

/* Operator coefficients alpha (curl-curl) and beta (mass).  If per_vector is
   non-null it points to device doubles (a.re, a.im, b.re, b.im) per vector. */
typedef struct {
  double ar, ai, br, bi;
  const double *per_vector;
} mxh_coef;

/* A block of nv vectors of one precision.  Vector v starts at
   p + (map ? map[v] : v) * stride complex values (map: device ints; only the
   launchers documented to honour it read it). */
typedef struct {
  void *p;
  long long stride; /* complex values between consecutive vectors */
  const int *map;
} mxh_vecs;

int mxh_init(void);
void *mxh_stream(void);
int mxh_sync(void);
const char *mxh_error_string(int err);

void *mxh_malloc(size_t bytes);
void mxh_free(void *p);
int mxh_h2d(void *dst, const void *src, size_t bytes);
int mxh_d2h(void *dst, const void *src, size_t bytes); /* synchronous */
/* Stream-ordered device-to-host copy that returns at once; dst must come from
   mxh_host_alloc (pinned) for the copy to overlap the host's work.  Wait for
   it with an event recorded after it. */
int mxh_d2h_async(void *dst, const void *src, size_t bytes);
void *mxh_host_alloc(size_t bytes);
void mxh_host_free(void *p);
int mxh_d2d(void *dst, const void *src, size_t bytes);
int mxh_memset(void *dst, int value, size_t bytes);

/* ---- edge space ---- */

/* y_v = (accumulate ? y_v : 0) + (alpha_v K + beta_v M) x_v */
int mxh_addeval(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                mxh_vecs x, mxh_vecs y, const void *eta, const mxh_coef *cf,
                int accumulate);

/* PINVIT defect: r_v = (K + (1 - lam_g) M) e_g for FP64 e (g = e.map[v] when
   mapped; the Bloch factors follow the same index), written in precision rprec
   at slot v of r; sums |r_v|^2 and |M e_v|^2 into out[2 v], out[2 v + 1]
   (device doubles).  lam: device doubles indexed by g (Ritz values of K+M). */
int mxh_defect(int rprec, const mxh_level *lv, const mxh_bloch *bl, int nv,
               mxh_vecs e, mxh_vecs r, const double *eta64, const double *lam,
               double *out, double *work);

/* One colour phase (colour = (j%2)*2 + (i%2) of the patch node) of the AFW
   vertex-patch smoother with operator alpha K + beta M. */
int mxh_afw_phase(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                  mxh_vecs b, mxh_vecs x, const void *eta, double alpha,
                  double beta, int colour);

/* One colour of the pointwise edge Gauss-Seidel smoother (one edge at a
   time) with operator alpha K + beta M: colours 0 / 1 are the x-edges of
   even / odd rows, 2 / 3 the y-edges of even / odd columns (nx, ny even). */
int mxh_edge_gs_phase(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                      mxh_vecs b, mxh_vecs x, const void *eta, const mxh_coef *cf,
                      int colour);

/* Tile-fused AFW sweeps with operator K + M (alpha = beta = 1), reading x
   and writing the result to xo (a different buffer: workgroups read halo
   nodes that other workgroups own).  mode bits:
   1 backward colour order (then nsweep must be 1), 2 x starts at zero,
   4 add P xc_in to x first (coarse correction of the V-cycle),
   8 after one forward sweep write bc = R (b - (K + M) x) and xc = 0 on the
   coarse level (vectors of the coarse level, stride bc.stride).
   Forward sweeps without mode 8 may be fused (nsweep > 1). */
int mxh_afw_tile(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                 mxh_vecs b, mxh_vecs x, mxh_vecs xo, const void *eta, int nsweep,
                 int mode, mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc);

/* Coarse b_v = R (bf_v - (K + M) xf_v); coarse x_v = 0. */
int mxh_resid_restrict(int prec, const mxh_level *fine, const mxh_bloch *bl,
                       int nv, mxh_vecs bf, mxh_vecs xf, const void *eta_f,
                       mxh_vecs bc, mxh_vecs xc);

/* f_v += alpha P c_v (alpha real) */
int mxh_prolong(int prec, const mxh_level *fine, const mxh_bloch *bl, int nv,
                mxh_vecs c, mxh_vecs f, double alpha);
/* c_v += alpha P^H f_v (alpha real) */
int mxh_restrict(int prec, const mxh_level *fine, const mxh_bloch *bl, int nv,
                 mxh_vecs f, mxh_vecs c, double alpha);

/* Per-problem coarse set-up on the device for an n0 x n0 coarse grid with
   2 n0^2 <= MXH_COARSE_SETUP_MAX: Ae[p] = chol(K + M) (2 n0^2 square) and
   An[p] = chol(node Laplacian, Gamma fill) (n0^2 square), complex,
   column-major, lower triangle (upper zeroed), from the Bloch factors
   bloch[4p..4p+3] (xf.re, xf.im, yf.re, yf.im) and the coarse 1/eps eta0.
   info[p] (device ints): 0, or the column of a non-positive pivot. */
#define MXH_COARSE_SETUP_MAX 64
int mxh_coarse_setup(int n0, int nprob, const double *bloch, const double *eta0, double *Ae,
                     double *An, int *info);

/* x_v = L^{-H} L^{-1} b_v for the dense Cholesky factor L (n x n complex
   double, column-major, device), n = vector length.  With Lper, vector v uses
   the factor of problem g / vpp, g = vmap ? vmap[v] : v. */
int mxh_coarse_solve(int prec, int n, const double *L, const double *Lper,
                     int vpp, const int *vmap, int nv, mxh_vecs b, mxh_vecs x);

/* Gram matrices of nprob blocks of m <= 48 FP64 vectors; block j is problem
   q = pmap ? pmap[j] : j, owning vectors q*m .. q*m+m-1 of u: for a >= b
   (pair p = a (a+1) / 2 + b), out[j][p] = (u_a^H K u_b, u_a^H M u_b) as two
   complex values, K = sum_cells eta c c^H without the 1/(hx hy) factor.
   out: device doubles, 4 m (m+1) / 2 per block.  work: device scratch of
   mxh_gram_work_bytes(m, nprob) bytes. */
size_t mxh_gram_work_bytes(int m, int nprob);
int mxh_gram(const mxh_level *lv, const mxh_bloch *bl, int nprob, const int *pmap,
             int m, mxh_vecs u, const double *eta64, double *out, void *work);

/* mxh_gram without the work argument: it allocates nothing (no hipMalloc,
   hipFree or allocating library call) and keeps its partial sums in a static
   device array of MXH_GRAM_SCRATCH_BYTES that the library holds on each
   device for the life of the process.  No counterpart in the reference: it is
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid allocation).
   Returns MXH_ESCRATCH, without launching, when nprob Gram sets of m vectors
   exceed the array (4 m (m+1) / 2 doubles each: nprob <= 3360 at m = 12,
   <= 873 at m = 24, <= 222 at m = 48).  The result equals mxh_gram's up to
   the summation order of the partial sums (bit for bit when m (m+1) / 2 <= 128).
   All calls on one device share the array: calls that overlap on one device
   (another host thread, another stream) race.  Needs mxh_init first. */
#define MXH_GRAM_SCRATCH_BYTES (8u << 20)
#define MXH_ESCRATCH (-2)
int mxh_gram_static(const mxh_level *lv, const mxh_bloch *bl, int nprob, const int *pmap,
                    int m, mxh_vecs u, const double *eta64, double *out);

/* mxh_gram over a subset of each block's columns: cmap (device ints, m + 1
   per problem q, indexed by q like the vectors) holds the count mq
   (1 <= mq <= m) and then mq increasing column indices below m.  Block j's
   sums run over those columns only and land at the pair indices of the list
   (pair a (a+1) / 2 + b of list positions a >= b, the same values mxh_gram
   gives for the columns at those positions); the entries past them are
   unspecified.  The LOBPCG Ritz step passes its block's E, P and unlocked W
   columns: a locked vector's W column is not written, so its products would
   be discarded. */
int mxh_gram_cols(const mxh_level *lv, const mxh_bloch *bl, int nprob, const int *pmap,
                  const int *cmap, int m, mxh_vecs u, const double *eta64, double *out,
                  void *work);

/* Per block j (problem q = pmap ? pmap[j] : j): E_j = sum_i C_j[i + j m] U_i
   over the problem's m FP64 vectors; C: device complex, m*m per block.
   n = complex values per vector. */
int mxh_block_combine(long long n, int nprob, const int *pmap, int m, mxh_vecs u,
                      mxh_vecs e, const double *C);

/* Per block j (problem q = pmap ? pmap[j] : j): for nout <= 32 outputs,
   out_k = sum_{a < nin} C_j[a + k nin] x_a over the problem's nin <= 48 FP64
   input vectors x_a = x.p + (q xvpp + a) x.stride, every input read once;
   out_k is written to y.p + (q yvpp + s_k) y.stride with s_k = k for
   k < nsplit and k + gap from there on.  C: device complex, nin*nout per block.
   n = complex values per vector. */
int mxh_combine_x(long long n, int nprob, const int *pmap, int nin, int nout, int nsplit, int gap,
                  mxh_vecs x, int xvpp, mxh_vecs y, int yvpp, const double *C);

/* LOBPCG block update per block j (problem q = pmap ? pmap[j] : j) of the
   basis X_q = [E | W | P] (3m FP64 vectors from x.p + 3 m q x.stride, m even,
   2 <= m <= 16) with the Ritz coefficients C_j (3m x 2m complex, column-major;
   only the first m columns, those of E_new, are read): P_new = [W P] C_WP
   and E_new = E C_E + P_new into the E and P slots of y's block q (W slots
   untouched). */
int mxh_lobpcg_combine(long long n, int nprob, const int *pmap, int m, mxh_vecs x, mxh_vecs y, const double *C);

/* dst_v = src_{map[v]} (compact gather, precision prec; src.map honoured). */
int mxh_gather(int prec, long long n, int nv, mxh_vecs dst, mxh_vecs src);

/* Householder M-orthonormalisation of k FP64 edge vectors (orthonormalize_edge2d,
   edge2d.c:6154-6296) entirely on the device: X holds the k inputs back to back
   (2 nx ny complex each) and receives the Householder vectors; V receives the
   M-orthonormal outputs V_i = H_0 ... H_i unit_i; tmp is one vector of
   scratch; tau receives the k complex tau_i (device); work is device scratch of
   mxh_orthonormalize_work_bytes(k).  Bloch factors xf = xr + i xi, yf = yr + i yi;
   eta64: 1/eps per cell.  k <= nx (ny / 2) unit vectors exist. */
size_t mxh_orthonormalize_work_bytes(int k);
int mxh_orthonormalize(const mxh_level *lv, double xr, double xi, double yr, double yi, const double *eta64,
                       int k, double *X, double *V, double *tmp, double *tau, double *work);

/* ---- node space ---- */
/* Tile-fused nodal Gauss-Seidel sweeps (alpha = 1), same modes as
   mxh_afw_tile; mode 8 writes bc = R (b - A x), xc = 0 on the coarse level. */
int mxh_node_tile(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                  mxh_vecs b, mxh_vecs x, mxh_vecs xo, int nsweep, int mode,
                  mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc);

/* Row-marching forms of mxh_afw_tile and mxh_node_tile: nsweep (1 or 2)
   sweeps per launch, the same modes (0, 1, 2, 5, 8, 10; zero start,
   prolongation at load, residual + restriction after the last sweep) and the
   same arithmetic, hence the same result.  A workgroup of 128 * nsweep
   threads owns up to 128 - 2H columns (halo H = 6 for one sweep, 10 for two)
   and `rows` rows (0: 128) and streams its region through an LDS ring of
   rows.  Requires nx, ny even and >= 10.  xo must differ from x (halo columns
   and rows are read from x by neighbours).  mxh_afw_march honours b.map
   (vector v's right-hand side is vector b.map[v] of b), so a caller can
   smooth a compacted set of vectors on right-hand sides left where they were
   formed. */
int mxh_afw_march(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                  mxh_vecs b, mxh_vecs x, mxh_vecs xo, const void *eta, int mode,
                  mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, int rows, int nsweep);
int mxh_node_march(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                   mxh_vecs b, mxh_vecs x, mxh_vecs xo, int mode,
                   mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc, int rows, int nsweep);
int mxh_node_addeval(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                     mxh_vecs x, mxh_vecs y, double ar, double ai, int accumulate);
int mxh_node_gs_phase(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                      mxh_vecs b, mxh_vecs x, double alpha, int colour);
int mxh_node_resid_restrict(int prec, const mxh_level *fine, const mxh_bloch *bl,
                            int nv, mxh_vecs bf, mxh_vecs xf, mxh_vecs bc,
                            mxh_vecs xc);
int mxh_node_prolong(int prec, const mxh_level *fine, const mxh_bloch *bl, int nv,
                     mxh_vecs c, mxh_vecs f, double ar, double ai);
int mxh_node_restrict(int prec, const mxh_level *fine, const mxh_bloch *bl, int nv,
                      mxh_vecs f, mxh_vecs c, double ar, double ai);

/* u_v = e_v - x_v (u, e FP64; x of precision prec, may be null); bg_v (prec)
   = G^H M u_v; xg_v = 0 (xg.p may be null: a caller whose node cycles start
   from zero by themselves skips the clearing).  u is written only where
   given (u.p non-null; with x null it is not written); it must not alias e
   when x is given.  e.map and u.map are honoured. */
/* With coef non-null (device, 2 doubles per vector; work: >= 2 * 513 * nv
   doubles) the lowest Bloch plane wave phi of the node space is deflated:
   bg -= (phi^H d / N) phi in FP64, coef_v = phi^H d / (N mu) with mu its
   Laplacian eigenvalue (zero at the Gamma point). */
int mxh_lift_rhs(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                 mxh_vecs e, mxh_vecs x, mxh_vecs u, mxh_vecs bg, mxh_vecs xg,
                 double *coef, double *work);
/* e_v (FP64) = u_v - x_v + alpha G (xg_v + a_v phi), a_v = coef_v -
   phi^H xg_v / N: the plane-wave component of xg (rounding only, after
   deflation) is replaced by the exact coef_v of mxh_lift_rhs.  x (precision
   prec, compact index v) may be null; coef may be null (plain G xg);
   otherwise work: >= 2 * 513 * nv doubles.  u.map and e.map are honoured; u
   may be e (in place); u.p null means u = 0 (e receives the correction). */
int mxh_lift_apply(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                   mxh_vecs xg, mxh_vecs u, mxh_vecs x, mxh_vecs e, double alpha,
                   const double *coef, double *work);

/* generic gradient / adjoint gradient in one precision */
int mxh_gradient(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                 mxh_vecs v, mxh_vecs e, double ar, double ai);
int mxh_adjgradient(int prec, const mxh_level *lv, const mxh_bloch *bl, int nv,
                    mxh_vecs e, mxh_vecs v, double ar, double ai);

/* ---- BLAS-1 (complex, n values per vector) ---- */
int mxh_scale(int prec, long long n, int nv, mxh_vecs x, double ar, double ai);
int mxh_axpy(int prec, long long n, int nv, mxh_vecs x, mxh_vecs y, double ar,
             double ai);
/* Reductions: work is device scratch of >= 512 * 4 * nv doubles.
   out[v] = x_v^H y_v (complex, device doubles 2 nv) */
int mxh_dot(int prec, long long n, int nv, mxh_vecs x, mxh_vecs y, double *out,
            double *work);
/* out[v] = |x_v|_2^2 or max |x_v| (kind 0/1) */
int mxh_norm(int prec, long long n, int nv, mxh_vecs x, int kind, double *out,
             double *work);
/* Gamma point: subtract the mean of the x-edge and of the y-edge values.
   work: >= 4 * nv * 513 doubles. */
int mxh_center(int prec, long long half, int nv, mxh_vecs x, double *work);
/* Convert n complex values between precisions. */
int mxh_convert(int dprec, int sprec, long long n, int nv, mxh_vecs dst,
                mxh_vecs src);
/* Seeded complex Gaussian entries, deterministic per (seed, vector, index). */
int mxh_random(int prec, long long n, int nv, mxh_vecs x, unsigned long long seed,
               int first_vector);

/* ---- adjoint ---- */
/* grad[c] += w_v * |curl e_v|_c^2 / (hx hy) summed over the nv vectors. */
int mxh_curl2(const mxh_level *lv, const mxh_bloch *bl, int nv, mxh_vecs e,
              const double *w, double *grad);

/* ---- events on the library stream (timing, and waiting for one point of
   the stream: mxh_event_sync) ---- */
int mxh_event_create(void **ev);
int mxh_event_record(void *ev);
int mxh_event_sync(void *ev);
int mxh_event_elapsed_ms(void *start, void *stop, float *ms);
int mxh_event_destroy(void *ev);

/* Workgroups of a row-marching launch (mxh_afw_march / mxh_node_march) with
   `rows` owned rows per workgroup and nsweep sweeps over nv vectors, and
   the compute units of the current device (for launch-shape choices). */
long long mxh_march_workgroups(const mxh_level *lv, int rows, int nsweep, int nv);
int mxh_cu_count(void);

/* ---- graphs: capture what is enqueued on the library stream ----
   begin; launches (no host-synchronous calls, no memsets); end, which
   instantiates into *exec, or updates *exec in place when it holds a graph
   of the same topology; launch replays it on the library stream.
   abort ends a capture that failed half-way and discards it. */
int mxh_capture_begin(void);
int mxh_capture_end(void **exec);
void mxh_capture_abort(void);
int mxh_graph_launch(void *exec);
void mxh_graph_destroy(void *exec);

#ifdef __cplusplus
}
#endif

#endif
