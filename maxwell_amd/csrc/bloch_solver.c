/*
 * bloch_solver.c -- batched block PINVIT with Rayleigh-Ritz on the GPU.
 *
 * One block step for every problem of the batch is the sequence of
 * exp_bloch2d.c:956-1053 (and exp_blochsurf.c:137-227):
 *   b = (K + (1 - lam) M) e                    defect (kept from the last step)
 *   x = V^{prec_steps}(b)                      edge V(nu,nu) cycles, x0 = 0
 *   u = e - x;  u -= G V_node^{5}(G^H M u)    gradient lifting (edge2d.c:6319-6327)
 *   [Gamma point: remove the constant fields, test_block2d.c:217-220]
 *   Ar = U^H (K+M) U, Mr = U^H M U; hegv; E = U C
 *   b = (K + (1 - lam) M) e and its norms     convergence test + next defect
 * Every kernel launch covers all vectors of all problems of the batch; the
 * small Ritz problems are solved on the host in FP64.
 */
#include "maxwell_amd.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mx_dense.h"
#include "mxh.h"

#define MX_MAXL 16
#define MX_TUNE_B 3 /* autotuned launch sizes: the full batch, 1/4 and 1/16 of it */

struct mx_solver {
  mx_config cfg;
  int L, n, prec;       /* levels, finest n, MG precision (MXH_F64 / MXH_F32) */
  size_t csz;           /* bytes per complex value of the MG precision */
  int nvmax;
  int lobpcg;           /* method: 0 block PINVIT, 1 LOBPCG */
  int vpp;              /* block-buffer vectors per problem: block (PINVIT) or
                           3 block (LOBPCG: the basis [E | W | P]) */
  mxh_level lv[MX_MAXL];
  double *eta64[MX_MAXL];
  void *etaP[MX_MAXL];
  double *hEta;         /* host eta per level (for the coarse assembly) */
  double *hEtaL[MX_MAXL];
  double *dE, *dU;      /* FP64 blocks at the finest level */
  void *eb[MX_MAXL], *ex[MX_MAXL];   /* edge MG vectors */
  void *nb[MX_MAXL], *nxv[MX_MAXL];  /* node MG vectors */
  void *ex2[MX_MAXL], *nxv2[MX_MAXL]; /* second x buffers: tile kernels write here, then swap */
  void *rdef;           /* defects of a block step (MG precision, defect slots) */
  double *dbloch, *dAe, *dAn, *dlam, *dnorm, *dgram, *dC, *dwork, *dcoef;
  void *gramwork;
  int ne0, nn0;         /* coarse edge / node unknowns */
  /* host scratch */
  double complex *hAe, *hAn, *hMr, *hC; /* hMr, hC: m x m per problem (PINVIT);
                                           hC: 3m x 2m per problem (LOBPCG) */
  int *hwvalid;         /* LOBPCG: vector's W column written this step */
  int pvalid;           /* LOBPCG: the P columns hold a search direction
                           (from the second Ritz step of the loop on) */
  int *hcols, *dcols;   /* LOBPCG: the Gram's column list per problem (vpp + 1
                           ints: count, then columns of [E W P]) */
  double *hgram, *hnorm, *hlam, *hbloch; /* hgram: pinned */
  void *ev_gram[2];                      /* Gram sums of each half on the host */
  int *gamma;
  int *hlist, *hslot;   /* host index lists (problems, vectors, maps) */
  int *dlist;           /* device copies of the lists */
  int *dinfo, *hinfo;   /* per-problem coarse factorisation status */
  int last_nprob;
  mx_stats st;
  /* smoother choice per level after autotuning (tuned = 1): 0 = square
     tiles, r > 0 = row-marching with r owned rows per workgroup */
  int tuned;
  int tune_nv[MX_TUNE_B];               /* vectors per launch each choice was timed for */
  int tune_rows[MX_TUNE_B][2][MX_MAXL]; /* [bucket][0: edge (AFW), 1: node][level] */
  /* the captured preconditioner of a block step (cfg.graphs), valid for
     graph_ng vectors, the block in graph_E and the multigrid buffer
     assignment graph_ptr */
  void *graph;
  int graph_ng;
  double *graph_E;
  void *graph_ptr[4][MX_MAXL];
  long long graph_launches, graph_edge, graph_node; /* counts of one replay */
};

#define CK(x)                      \
  do {                             \
    int _e = (x);                  \
    if (_e) return MX_EDEVICE;     \
  } while (0)

static mxh_vecs V(void *p, long long stride)
{
  mxh_vecs v;
  v.p = p;
  v.stride = stride;
  v.map = NULL;
  return v;
}

static long long edim(const mx_solver *s, int l) { return 2LL * s->lv[l].nx * s->lv[l].ny; }
static long long ndim(const mx_solver *s, int l) { return (long long)s->lv[l].nx * s->lv[l].ny; }

void mx_solver_destroy(mx_solver *s)
{
  if (!s) return;
  mxh_graph_destroy(s->graph);
  for (int l = 0; l <= s->L; l++) {
    mxh_free(s->eta64[l]);
    if (s->etaP[l] != s->eta64[l]) mxh_free(s->etaP[l]);
    mxh_free(s->eb[l]);
    mxh_free(s->ex[l]);
    mxh_free(s->nb[l]);
    mxh_free(s->nxv[l]);
    mxh_free(s->ex2[l]);
    mxh_free(s->nxv2[l]);
    free(s->hEtaL[l]);
  }
  mxh_free(s->rdef);
  mxh_free(s->dE);
  mxh_free(s->dU);
  mxh_free(s->dbloch);
  mxh_free(s->dAe);
  mxh_free(s->dAn);
  mxh_free(s->dlam);
  mxh_free(s->dnorm);
  mxh_free(s->dgram);
  mxh_free(s->dC);
  mxh_free(s->dwork);
  mxh_free(s->dcoef);
  mxh_free(s->gramwork);
  free(s->hAe);
  free(s->hAn);
  free(s->hMr);
  free(s->hC);
  free(s->hwvalid);
  mxh_host_free(s->hgram);
  for (int h = 0; h < 2; h++)
    if (s->ev_gram[h]) mxh_event_destroy(s->ev_gram[h]);
  free(s->hnorm);
  free(s->hlam);
  free(s->hbloch);
  free(s->gamma);
  free(s->hlist);
  free(s->hslot);
  mxh_free(s->dlist);
  free(s->hcols);
  mxh_free(s->dcols);
  mxh_free(s->dinfo);
  free(s->hinfo);
  free(s);
}

/* device allocation counted into the solver's footprint */
static void *dalloc(mx_solver *s, size_t bytes)
{
  void *p = mxh_malloc(bytes);
  if (p) s->st.device_bytes += (long long)bytes;
  return p;
}

int mx_solver_create(const mx_config *cfg, const double *eta, mx_solver **out)
{
  *out = NULL;
  if (!cfg || !eta || cfg->n0 < 2 || (cfg->n0 & 1) || cfg->levels < 0 || cfg->levels >= MX_MAXL ||
      cfg->block < 1 || cfg->block > 32 || cfg->max_problems < 1 || cfg->smoother_steps < 0 ||
      cfg->prec_steps < 1 || cfg->gradient_steps < 1 || cfg->march_min < 0 || cfg->march_rows < 0 ||
      cfg->march_fuse < 0 || cfg->march_fuse > 1 || cfg->autotune < 0 || cfg->autotune > 1 ||
      cfg->graphs < 0 || cfg->graphs > 1 || (cfg->precision != MX_PREC_F64 && cfg->precision != MX_PREC_MIXED) ||
      (cfg->method != MX_METHOD_PINVIT && cfg->method != MX_METHOD_LOBPCG) ||
      (cfg->method == MX_METHOD_LOBPCG && (3 * cfg->block > 48 || !cfg->lock)))
    return MX_EINVAL;
  if (2 * cfg->n0 * cfg->n0 > 256) return MX_EINVAL; /* coarse solve: one block */
  {
    /* 1/eps must be finite and positive in every cell: K is then positive
       semi-definite with the gradients as its null space */
    const size_t nc = (size_t)(cfg->n0 << cfg->levels) * (size_t)(cfg->n0 << cfg->levels);
    for (size_t c = 0; c < nc; c++)
      if (!(eta[c] > 0.0) || !isfinite(eta[c])) return MX_EINVAL;
  }
  if (mxh_init()) return MX_EDEVICE;
  mx_solver *s = (mx_solver *)calloc(1, sizeof(mx_solver));
  if (!s) return MX_ENOMEM;
  s->cfg = *cfg;
  s->L = cfg->levels;
  s->n = cfg->n0 << cfg->levels;
  s->prec = cfg->precision == MX_PREC_MIXED ? MXH_F32 : MXH_F64;
  s->csz = s->prec == MXH_F64 ? 16 : 8;
  s->nvmax = cfg->max_problems * cfg->block;
  s->lobpcg = cfg->method == MX_METHOD_LOBPCG;
  s->vpp = s->lobpcg ? 3 * cfg->block : cfg->block;
  s->ne0 = 2 * cfg->n0 * cfg->n0;
  s->nn0 = cfg->n0 * cfg->n0;
  const int P = cfg->max_problems, m = cfg->block, nv = s->nvmax, vpp = s->vpp;

  for (int l = 0; l <= s->L; l++) {
    int nl = cfg->n0 << l;
    s->lv[l].nx = nl;
    s->lv[l].ny = nl;
    s->lv[l].hx = 1.0 / nl;
    s->lv[l].hy = 1.0 / nl;
    s->hEtaL[l] = (double *)malloc(sizeof(double) * (size_t)nl * nl);
    if (!s->hEtaL[l]) goto nomem;
  }
  memcpy(s->hEtaL[s->L], eta, sizeof(double) * (size_t)s->n * s->n);
  for (int l = s->L; l > 0; l--) { /* grid2d.c:48-68 */
    int nf = s->lv[l].nx, nc = nf / 2;
    const double *ef = s->hEtaL[l];
    double *ec = s->hEtaL[l - 1];
    for (int j = 0; j < nc; j++)
      for (int i = 0; i < nc; i++)
        ec[i + j * nc] = 0.25 * (ef[2 * i + 2 * j * nf] + ef[2 * i + 1 + 2 * j * nf] +
                                 ef[2 * i + (2 * j + 1) * nf] + ef[2 * i + 1 + (2 * j + 1) * nf]);
  }
  for (int l = 0; l <= s->L; l++) {
    size_t N = (size_t)ndim(s, l);
    s->eta64[l] = (double *)dalloc(s, sizeof(double) * N);
    if (!s->eta64[l]) goto nomem;
    if (mxh_h2d(s->eta64[l], s->hEtaL[l], sizeof(double) * N)) goto dev;
    if (s->prec == MXH_F64) {
      s->etaP[l] = s->eta64[l];
    } else {
      float *tmp = (float *)malloc(sizeof(float) * N);
      if (!tmp) goto nomem;
      for (size_t k = 0; k < N; k++) tmp[k] = (float)s->hEtaL[l][k];
      s->etaP[l] = dalloc(s, sizeof(float) * N);
      int e = s->etaP[l] ? mxh_h2d(s->etaP[l], tmp, sizeof(float) * N) : 1;
      if (!e) e = mxh_sync();
      free(tmp);
      if (!s->etaP[l]) goto nomem;
      if (e) goto dev;
    }
    s->eb[l] = dalloc(s, s->csz * (size_t)edim(s, l) * nv);
    s->ex[l] = dalloc(s, s->csz * (size_t)edim(s, l) * nv);
    s->nb[l] = dalloc(s, s->csz * (size_t)ndim(s, l) * nv);
    s->nxv[l] = dalloc(s, s->csz * (size_t)ndim(s, l) * nv);
    s->ex2[l] = dalloc(s, s->csz * (size_t)edim(s, l) * nv);
    s->nxv2[l] = dalloc(s, s->csz * (size_t)ndim(s, l) * nv);
    if (!s->eb[l] || !s->ex[l] || !s->nb[l] || !s->nxv[l] || !s->ex2[l] || !s->nxv2[l]) goto nomem;
  }
  s->rdef = dalloc(s, s->csz * (size_t)edim(s, s->L) * nv);
  if (!s->rdef) goto nomem;
  s->dE = (double *)dalloc(s, 16 * (size_t)edim(s, s->L) * vpp * P);
  s->dU = (double *)dalloc(s, 16 * (size_t)edim(s, s->L) * vpp * P);
  s->dbloch = (double *)dalloc(s, sizeof(double) * 4 * P);
  s->dAe = (double *)dalloc(s, 16 * (size_t)s->ne0 * s->ne0 * P);
  s->dAn = (double *)dalloc(s, 16 * (size_t)s->nn0 * s->nn0 * P);
  s->dlam = (double *)dalloc(s, sizeof(double) * (size_t)vpp * P);
  s->dnorm = (double *)dalloc(s, sizeof(double) * 2 * nv);
  s->dgram = (double *)dalloc(s, sizeof(double) * 4 * (size_t)vpp * vpp * P);
  s->dC = (double *)dalloc(s, 16 * (size_t)vpp * m * 2 * P);
  s->dwork = (double *)dalloc(s, sizeof(double) * (513 * 4 * (size_t)nv + 1024));
  s->dcoef = (double *)dalloc(s, sizeof(double) * 2 * (size_t)nv);
  s->gramwork = dalloc(s, mxh_gram_work_bytes(vpp, P));
  if (!s->dE || !s->dU || !s->dbloch || !s->dAe || !s->dAn || !s->dlam || !s->dnorm || !s->dgram || !s->dC ||
      !s->dwork || !s->dcoef || !s->gramwork)
    goto nomem;
  s->hAe = (double complex *)malloc(sizeof(double complex) * (size_t)s->ne0 * s->ne0 * P);
  s->hAn = (double complex *)malloc(sizeof(double complex) * (size_t)s->nn0 * s->nn0 * P);
  s->hMr = (double complex *)malloc(sizeof(double complex) * (size_t)m * m * P);
  s->hC = (double complex *)malloc(sizeof(double complex) * (size_t)vpp * m * 2 * P);
  s->hwvalid = (int *)calloc((size_t)nv, sizeof(int));
  s->hgram = (double *)mxh_host_alloc(sizeof(double) * 4 * (size_t)vpp * vpp * P);
  for (int h = 0; h < 2; h++)
    if (mxh_event_create(&s->ev_gram[h])) s->ev_gram[h] = NULL;
  s->hnorm = (double *)malloc(sizeof(double) * 2 * nv);
  s->hlam = (double *)malloc(sizeof(double) * (size_t)vpp * P);
  s->hbloch = (double *)malloc(sizeof(double) * 4 * P);
  s->gamma = (int *)calloc(P, sizeof(int));
  s->hlist = (int *)malloc(sizeof(int) * (P + 4 * (size_t)nv));
  s->hslot = (int *)malloc(sizeof(int) * (size_t)vpp * P);
  s->dlist = (int *)dalloc(s, sizeof(int) * (P + 4 * (size_t)nv));
  s->hcols = (int *)malloc(sizeof(int) * (size_t)(vpp + 1) * P);
  s->dcols = (int *)dalloc(s, sizeof(int) * (size_t)(vpp + 1) * P);
  s->dinfo = (int *)dalloc(s, sizeof(int) * P);
  s->hinfo = (int *)malloc(sizeof(int) * P);
  if (!s->hAe || !s->hAn || !s->hMr || !s->hC || !s->hwvalid || !s->hgram || !s->ev_gram[0] || !s->ev_gram[1] || !s->hnorm || !s->hlam || !s->hbloch ||
      !s->gamma || !s->hlist || !s->hslot || !s->dlist || !s->dinfo || !s->hinfo || !s->hcols || !s->dcols)
    goto nomem;
  if (mxh_sync()) goto dev;
  *out = s;
  return MX_OK;
nomem:
  mx_solver_destroy(s);
  return MX_ENOMEM;
dev:
  mx_solver_destroy(s);
  return MX_EDEVICE;
}

static mxh_bloch batch_bloch(const mx_solver *s)
{
  mxh_bloch b;
  memset(&b, 0, sizeof(b));
  b.xr = 1.0;
  b.yr = 1.0;
  b.per_problem = s->dbloch;
  b.vpp = s->vpp;
  return b;
}

static void swap_ptr(void **a, void **b)
{
  void *t = *a;
  *a = *b;
  *b = t;
}

/* ns smoothing sweeps (1, or 2 in one pass) on level l, ex[l] -> ex2[l]
   (nxv[l] -> nxv2[l]): the row-marching kernel on levels of at least
   march_min intervals, the square tile below. */
/* the autotuned launch size nearest to nv (ratio) */
static int tune_bucket(const mx_solver *s, int nv)
{
  int best = 0;
  double bd = 1e300;
  for (int b = 0; b < MX_TUNE_B; b++) {
    if (s->tune_nv[b] < 1) continue;
    const double d = fabs(log((double)nv / s->tune_nv[b]));
    if (d < bd) bd = d, best = b;
  }
  return best;
}

/* rows per marching workgroup on level l of smoother `node` for a launch
   over nv vectors (0: square tiles): the autotuned choice for the nearest
   tuned launch size, else march_rows on levels >= march_min */
static int march_rows_of(const mx_solver *s, int node, int l, int nv)
{
  if (s->lv[l].nx < 8 || s->lv[l].ny < 8) return 0; /* halo of one sweep: 6 */
  if (s->tuned) return s->tune_rows[tune_bucket(s, nv)][node][l];
  if (s->cfg.march_min <= 0) return 0;
  const int rows = s->cfg.march_rows > 0 ? s->cfg.march_rows : 128;
  if (s->lv[l].nx >= s->cfg.march_min) return rows;
  /* one level below march_min the marching tiles still win when the batch
     gives every compute unit three of them; with fewer vectors (locking)
     the square tiles' finer decomposition keeps the device busier */
  if (2 * s->lv[l].nx >= s->cfg.march_min && mxh_march_workgroups(&s->lv[l], rows, 2, nv) >= 3LL * mxh_cu_count())
    return rows;
  return 0;
}

/* sweeps to run in the next launch on level l when k of nu are done: two
   when fusing, for marching passes (grids of at least 10 x 10) and square
   tiles alike */
static int sweeps_next(const mx_solver *s, int node, int l, int nv, int k, int nu)
{
  if (!s->cfg.march_fuse || k + 2 > nu) return 1;
  if (march_rows_of(s, node, l, nv) > 0) return s->lv[l].nx >= 10 && s->lv[l].ny >= 10 ? 2 : 1;
  return 2;
}

static int afw_sweep(mx_solver *s, int l, int nv, const mxh_bloch *bl, mxh_vecs b, int mode, int ns, mxh_vecs xc_in,
                     mxh_vecs bc, mxh_vecs xc)
{
  mxh_vecs x = V(s->ex[l], edim(s, l)), xo = V(s->ex2[l], edim(s, l));
  const int rows = march_rows_of(s, 0, l, nv);
  if (rows > 0)
    return mxh_afw_march(s->prec, &s->lv[l], bl, nv, b, x, xo, s->etaP[l], mode, xc_in, bc, xc, rows, ns);
  return mxh_afw_tile(s->prec, &s->lv[l], bl, nv, b, x, xo, s->etaP[l], ns, mode, xc_in, bc, xc);
}

static int node_sweep(mx_solver *s, int l, int nv, const mxh_bloch *bl, mxh_vecs b, int mode, int ns,
                      mxh_vecs xc_in, mxh_vecs bc, mxh_vecs xc)
{
  mxh_vecs x = V(s->nxv[l], ndim(s, l)), xo = V(s->nxv2[l], ndim(s, l));
  const int rows = march_rows_of(s, 1, l, nv);
  if (rows > 0) return mxh_node_march(s->prec, &s->lv[l], bl, nv, b, x, xo, mode, xc_in, bc, xc, rows, ns);
  return mxh_node_tile(s->prec, &s->lv[l], bl, nv, b, x, xo, ns, mode, xc_in, bc, xc);
}

/* Autotuner: for every level, both smoothers and three launch sizes (the
   batch's nv vectors, nv/4 and nv/16: locking shrinks the launches as
   problems converge), time nu smoothing sweeps with the square tiles and
   with row-marching tiles of 32..256 rows (pairs of sweeps fused when
   march_fuse), HIP events on the library stream, on random data in the
   level's multigrid buffers (which the solve overwrites), and keep the
   fastest.  A launch uses the choice of the nearest tuned size. */
static int autotune(mx_solver *s, int nv_full, const mxh_bloch *bl)
{
  static const int cand[] = {0, 32, 64, 128, 192, 256};
  const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
  const int nu = s->cfg.smoother_steps > 0 ? s->cfg.smoother_steps : 1;
  void *e0, *e1;
  CK(mxh_event_create(&e0));
  if (mxh_event_create(&e1)) {
    mxh_event_destroy(e0);
    return MX_EDEVICE;
  }
  int err = 0;
  s->tuned = 0;
  mxh_graph_destroy(s->graph); /* captured with the previous tiling */
  s->graph = NULL;
  for (int b = 0; b < MX_TUNE_B; b++) s->tune_nv[b] = 0;
  for (int bk = 0; bk < MX_TUNE_B && !err; bk++) {
    const int nv = nv_full >> (2 * bk);
    if (nv < 1 || (bk > 0 && nv == s->tune_nv[bk - 1])) break;
    s->tune_nv[bk] = nv;
    for (int node = 0; node < 2 && !err; node++)
      for (int l = 1; l <= s->L && !err; l++) {
        const long long n = node ? ndim(s, l) : edim(s, l);
        void **bb = node ? s->nb : s->eb, **xx = node ? s->nxv : s->ex;
        err |= mxh_random(s->prec, n, nv, V(bb[l], n), 7 + l, 0);
        err |= mxh_random(s->prec, n, nv, V(xx[l], n), 11 + l, 0);
        float best = 1e30f;
        int pick = 0;
        for (int ci = 0; ci < ncand && !err; ci++) {
          const int rows = cand[ci];
          if (rows > 0 && (s->lv[l].nx < 8 || s->lv[l].ny < 8)) continue;
          if (rows > 2 * s->lv[l].ny) continue; /* same tiling as a smaller candidate */
          s->tuned = 1;
          s->tune_rows[bk][node][l] = rows;
          float ms = 0.0f;
          for (int rep = 0; rep < 3 && !err; rep++) { /* rep 0 warms up */
            err |= mxh_event_record(e0);
            for (int k = 0; k < nu && !err;) {
              const int ns = sweeps_next(s, node, l, nv, k, nu);
              mxh_vecs b = V(bb[l], n);
              err |= node ? node_sweep(s, l, nv, bl, b, 0, ns, V(NULL, 0), V(NULL, 0), V(NULL, 0))
                          : afw_sweep(s, l, nv, bl, b, 0, ns, V(NULL, 0), V(NULL, 0), V(NULL, 0));
              if (node) swap_ptr(&s->nxv[l], &s->nxv2[l]);
              else swap_ptr(&s->ex[l], &s->ex2[l]);
              k += ns;
            }
            err |= mxh_event_record(e1);
            float t = 0.0f;
            err |= mxh_event_elapsed_ms(e0, e1, &t);
            if (rep > 0) ms += t;
          }
          if (!err && ms < best) {
            best = ms;
            pick = rows;
          }
        }
        s->tune_rows[bk][node][l] = pick;
      }
  }
  mxh_event_destroy(e0);
  mxh_event_destroy(e1);
  s->tuned = !err;
  return err ? MX_EDEVICE : MX_OK;
}

/* V(nu,nu) on the edge hierarchy for nv vectors: eb[L] rhs (or bfine when
   given: the finest level's right-hand sides, read through its map by the
   row-marching smoother), ex[L] in/out.  xzero: ex[L] is zero on entry (it
   is then not read). */
static int vcycle_edge(mx_solver *s, int nv, const mxh_bloch *bl, int xzero, mxh_vecs bfine)
{
  const int P = s->prec, nu = s->cfg.smoother_steps;
  const mxh_vecs none = V(NULL, 0);
  for (int l = s->L; l > 0; l--) {
    mxh_vecs b = l == s->L && bfine.p ? bfine : V(s->eb[l], edim(s, l)), x = V(s->ex[l], edim(s, l));
    mxh_vecs bc = V(s->eb[l - 1], edim(s, l - 1)), xc = V(s->ex[l - 1], edim(s, l - 1));
    const int xz = (l < s->L) || xzero;
    if (s->cfg.phase_kernels || nu == 0) {
      if (xz && l == s->L) CK(mxh_memset(s->ex[l], 0, s->csz * (size_t)edim(s, l) * nv));
      for (int k = 0; k < nu; k++)
        for (int c = 0; c < 4; c++) CK(mxh_afw_phase(P, &s->lv[l], bl, nv, b, x, s->etaP[l], 1.0, 1.0, c));
      CK(mxh_resid_restrict(P, &s->lv[l], bl, nv, b, x, s->etaP[l], bc, xc));
      s->st.launches += 4 * nu + 1;
    } else {
      for (int k = 0; k < nu;) {
        const int ns = sweeps_next(s, 0, l, nv, k, nu);
        const int mode = (k == 0 && xz ? 2 : 0) | (k + ns == nu ? 8 : 0);
        CK(afw_sweep(s, l, nv, bl, b, mode, ns, none, bc, xc));
        swap_ptr(&s->ex[l], &s->ex2[l]);
        s->st.launches++;
        k += ns;
      }
    }
  }
  CK(mxh_coarse_solve(P, s->ne0, NULL, s->dAe, s->vpp, bl->vmap, nv, V(s->eb[0], edim(s, 0)),
                      V(s->ex[0], edim(s, 0))));
  s->st.launches += 1;
  for (int l = 1; l <= s->L; l++) {
    mxh_vecs b = l == s->L && bfine.p ? bfine : V(s->eb[l], edim(s, l)), x = V(s->ex[l], edim(s, l));
    mxh_vecs xc = V(s->ex[l - 1], edim(s, l - 1));
    if (s->cfg.phase_kernels || nu == 0) {
      CK(mxh_prolong(P, &s->lv[l], bl, nv, xc, x, 1.0));
      for (int k = 0; k < nu; k++)
        for (int c = 3; c >= 0; c--) CK(mxh_afw_phase(P, &s->lv[l], bl, nv, b, x, s->etaP[l], 1.0, 1.0, c));
      s->st.launches += 4 * nu + 1;
    } else {
      for (int k = 0; k < nu;) {
        const int ns = sweeps_next(s, 0, l, nv, k, nu);
        CK(afw_sweep(s, l, nv, bl, b, k == 0 ? 5 : 1, ns, xc, none, none));
        swap_ptr(&s->ex[l], &s->ex2[l]);
        s->st.launches++;
        k += ns;
      }
    }
  }
  s->st.edge_vcycles++;
  return MX_OK;
}

static int vcycle_node(mx_solver *s, int nv, const mxh_bloch *bl, int xzero)
{
  const int P = s->prec, nu = s->cfg.smoother_steps;
  const mxh_vecs none = V(NULL, 0);
  for (int l = s->L; l > 0; l--) {
    mxh_vecs b = V(s->nb[l], ndim(s, l)), x = V(s->nxv[l], ndim(s, l));
    mxh_vecs bc = V(s->nb[l - 1], ndim(s, l - 1)), xc = V(s->nxv[l - 1], ndim(s, l - 1));
    const int xz = (l < s->L) || xzero;
    if (s->cfg.phase_kernels || nu == 0) {
      if (xz && l == s->L) CK(mxh_memset(s->nxv[l], 0, s->csz * (size_t)ndim(s, l) * nv));
      for (int k = 0; k < nu; k++)
        for (int c = 0; c < 4; c++) CK(mxh_node_gs_phase(P, &s->lv[l], bl, nv, b, x, 1.0, c));
      CK(mxh_node_resid_restrict(P, &s->lv[l], bl, nv, b, x, bc, xc));
      s->st.launches += 4 * nu + 1;
    } else {
      for (int k = 0; k < nu;) {
        const int ns = sweeps_next(s, 1, l, nv, k, nu);
        const int mode = (k == 0 && xz ? 2 : 0) | (k + ns == nu ? 8 : 0);
        CK(node_sweep(s, l, nv, bl, b, mode, ns, none, bc, xc));
        swap_ptr(&s->nxv[l], &s->nxv2[l]);
        s->st.launches++;
        k += ns;
      }
    }
  }
  CK(mxh_coarse_solve(P, s->nn0, NULL, s->dAn, s->vpp, bl->vmap, nv, V(s->nb[0], ndim(s, 0)),
                      V(s->nxv[0], ndim(s, 0))));
  s->st.launches += 1;
  for (int l = 1; l <= s->L; l++) {
    mxh_vecs b = V(s->nb[l], ndim(s, l)), x = V(s->nxv[l], ndim(s, l));
    mxh_vecs xc = V(s->nxv[l - 1], ndim(s, l - 1));
    if (s->cfg.phase_kernels || nu == 0) {
      CK(mxh_node_prolong(P, &s->lv[l], bl, nv, xc, x, 1.0, 0.0));
      for (int k = 0; k < nu; k++)
        for (int c = 3; c >= 0; c--) CK(mxh_node_gs_phase(P, &s->lv[l], bl, nv, b, x, 1.0, c));
      s->st.launches += 4 * nu + 1;
    } else {
      for (int k = 0; k < nu;) {
        const int ns = sweeps_next(s, 1, l, nv, k, nu);
        CK(node_sweep(s, l, nv, bl, b, k == 0 ? 5 : 1, ns, xc, none, none));
        swap_ptr(&s->nxv[l], &s->nxv2[l]);
        s->st.launches++;
        k += ns;
      }
    }
  }
  s->st.node_vcycles++;
  return MX_OK;
}

/* Centre the vectors of the listed problems that sit at the Gamma point. */
static int center_gamma(mx_solver *s, int na, const int *plist, double *E)
{
  const int m = s->cfg.block;
  const long long n = edim(s, s->L);
  for (int j = 0; j < na; j++) {
    const int p = plist ? plist[j] : j;
    if (s->gamma[p]) {
      CK(mxh_center(MXH_F64, n / 2, m, V(E + 2 * (size_t)p * s->vpp * n, n), s->dwork));
      s->st.launches += 3;
    }
  }
  return MX_OK;
}

/* For the vectors g = gmap[a] (or all when gmap is null), with u_g = e_g - x_a
   (x in the MG precision, compact index a; none: u = e):
   e_g = u_g - G V_node^k (G^H M u_g), in place in E (u itself is never
   stored: both kernels form it from e and x); the other vectors of E are
   left as they are.  With W non-null the result minus e_g goes to vector
   wmap[a] of W instead (LOBPCG's search direction), and E is not written. */
static int lift_to(mx_solver *s, int nv, const mxh_bloch *bl, double *E, const void *X, const int *gmap, double *W,
                   const int *wmap)
{
  const int L = s->L;
  const long long n = edim(s, L);
  mxh_vecs ev = V(E, n), xv = V((void *)X, n);
  ev.map = gmap;
  /* the lowest Bloch plane wave of the node space (Laplacian eigenvalue
     ~(k/n)^2 near Gamma, below FP32 resolution and near the FP64 floor of an
     iterative solve) is handled exactly in FP64 in both precisions */
  double *coef = s->dcoef;
  /* the first node V-cycle (gradient_steps >= 1) starts from zero without
     reading xg (its smoothers' zero-start modes, or a clear of its own), so
     xg is not cleared here */
  CK(mxh_lift_rhs(s->prec, &s->lv[L], bl, nv, ev, xv, V(NULL, n), V(s->nb[L], ndim(s, L)), V(NULL, ndim(s, L)),
                  coef, s->dwork));
  s->st.launches += coef ? 3 : 1;  /* rhs, partial-sum reduction, deflation */
  for (int k = 0; k < s->cfg.gradient_steps; k++) {
    int e = vcycle_node(s, nv, bl, k == 0);
    if (e) return e;
  }
  mxh_vecs wv = V(W, n);
  wv.map = wmap;
  CK(mxh_lift_apply(s->prec, &s->lv[L], bl, nv, V(s->nxv[L], ndim(s, L)), W ? V(NULL, n) : ev, xv, W ? wv : ev,
                    -1.0, coef, s->dwork));
  s->st.launches += coef ? 3 : 1;
  return MX_OK;
}

static int lift(mx_solver *s, int nv, const mxh_bloch *bl, double *E, const void *X, const int *gmap)
{
  return lift_to(s, nv, bl, E, X, gmap, NULL, NULL);
}

/* The small generalised eigenproblems Ar c = lam Mr c (Ar = U^H (K+M) U,
   Mr = U^H M U, from the Gram sums in hgram) of the list's problems
   j0 <= j < j1: eigenvectors to hC, Ritz values (of K+M) to hlam. */
static int ritz_host(mx_solver *s, int j0, int j1, const int *plist)
{
  const int m = s->cfg.block, L = s->L, np = m * (m + 1) / 2;
  const double sk = 1.0 / (s->lv[L].hx * s->lv[L].hy);
  for (int j = j0; j < j1; j++) {
    const int p = plist ? plist[j] : j;
    const double *g = s->hgram + 4 * (size_t)np * j;
    double complex *Ar = s->hC + (size_t)j * m * m, *Mr = s->hMr + (size_t)j * m * m;
    for (int q = 0; q < 4 * np; q++)
      if (!isfinite(g[q])) return MX_ENOTFINITE;
    for (int a = 0; a < m; a++)
      for (int c = 0; c <= a; c++) {
        const double *q = g + 4 * (a * (a + 1) / 2 + c);
        double complex gk = q[0] + I * q[1], gm = q[2] + I * q[3];
        Ar[a + c * m] = sk * gk + gm;
        Mr[a + c * m] = gm;
        Ar[c + a * m] = conj(Ar[a + c * m]);
        Mr[c + a * m] = conj(gm);
      }
    /* Ar becomes the eigenvectors C of the problem.  Mr must be definite
       to well above rounding: a pivot of its Cholesky factor below 1e-13 of
       its diagonal entry means block vectors that are dependent to FP64
       precision (repeated or zero start vectors), whose Ritz values would be
       rounding noise */
    if (mxd_hegv_rel(m, Ar, Mr, s->hlam + (size_t)p * s->vpp, 1e-13)) return MX_ELAPACK;
  }
  return MX_OK;
}

/* LOBPCG Rayleigh-Ritz of the list's problems j0 <= j < j1 over the basis
   [E | W | P] (3m columns; the Gram sums of all pairs in hgram): the columns
   are scaled to unit M-norm and taken in the order E, W, P, each kept only
   if the squared sine of its angle to the span of the columns kept before it
   exceeds LOBPCG_PIVOT (pivot of a Cholesky factorisation of the scaled
   M-Gram), so that W and P columns that are zero (no step yet, or a locked
   vector's W, which is not written while the vector is locked) or
   numerically dependent drop out.  The kept basis then has a scaled Gram of
   condition below about 3m / LOBPCG_PIVOT, which bounds what the rounding of
   the Gram sums (eps relative) does to the M-orthonormality of the new block
   (about eps times that condition).  At least m columns must stay.  The
   lowest m Ritz pairs of the kept pencil give hlam and two coefficient blocks
   per problem in hC (3m x 2m, column-major): E_new = [E W P] C_E and the next
   search direction P_new = [W P] C_WP (the W and P rows of C_E). */
#define LOBPCG_PIVOT 1e-7
static int lobpcg_problem(mx_solver *s, int j, int p)
{
  const int m = s->cfg.block, L = s->L, k = 3 * m, np = k * (k + 1) / 2;
  const double sk = 1.0 / (s->lv[L].hx * s->lv[L].hy);
  double w[48], dsc[48];
  int keep[48];
  double complex A[48 * 48], B[48 * 48], Lf[48 * 48], As[48 * 48], Bs[48 * 48];
  const double *g = s->hgram + 4 * (size_t)np * j;
  /* the Gram holds the problem's listed columns (E, the W of its unlocked
     vectors, P once it holds a direction) at list positions 0 .. kc - 1 */
  const int *cl = s->hcols + (size_t)p * (k + 1), kc = cl[0];
  for (int q = 0; q < 2 * kc * (kc + 1); q++)
    if (!isfinite(g[q])) return MX_ENOTFINITE;
  for (int a = 0; a < kc; a++)
    for (int c = 0; c <= a; c++) {
      const double *q = g + 4 * (a * (a + 1) / 2 + c);
      const double complex gk = q[0] + I * q[1], gm = q[2] + I * q[3];
      A[a + c * k] = sk * gk + gm;
      B[a + c * k] = gm;
      A[c + a * k] = conj(A[a + c * k]);
      B[c + a * k] = conj(gm);
    }
  /* greedy column selection by a Cholesky factorisation of the scaled B */
  int kk = 0;
  for (int c = 0; c < kc; c++) {
    const double bcc = creal(B[c + c * k]);
    if (!(bcc > 0.0) || !isfinite(bcc)) continue;
    const double d = 1.0 / sqrt(bcc);
    /* row kk of the factor of the scaled B restricted to the kept columns */
    double piv = 1.0;
    for (int t = 0; t < kk; t++) {
      double complex z = B[c + keep[t] * k] * d * dsc[t];
      for (int u = 0; u < t; u++) z -= Lf[kk + u * 48] * conj(Lf[t + u * 48]);
      z /= creal(Lf[t + t * 48]);
      Lf[kk + t * 48] = z;
      piv -= creal(z * conj(z));
    }
    if (!(piv > LOBPCG_PIVOT)) continue;
    Lf[kk + kk * 48] = sqrt(piv);
    keep[kk] = c;
    dsc[kk] = d;
    kk++;
  }
  if (kk < m) return MX_ELAPACK;
  for (int t = 0; t < kk; t++)
    for (int u = 0; u < kk; u++) {
      As[t + u * kk] = A[keep[t] + keep[u] * k] * dsc[t] * dsc[u];
      Bs[t + u * kk] = B[keep[t] + keep[u] * k] * dsc[t] * dsc[u];
    }
  if (mxd_hegv(kk, As, Bs, w)) return MX_ELAPACK;
  double complex *Cp = s->hC + (size_t)j * k * 2 * m;
  for (int q = 0; q < k * 2 * m; q++) Cp[q] = 0.0;
  for (int i = 0; i < m; i++) {
    s->hlam[(size_t)p * s->vpp + i] = w[i];
    for (int t = 0; t < kk; t++) {
      const int c = cl[1 + keep[t]];
      const double complex cf = As[t + i * kk] * dsc[t];
      Cp[c + (size_t)i * k] = cf;
      if (c >= m) Cp[c + (size_t)(m + i) * k] = cf;
    }
  }
  return MX_OK;
}

/* The problems are independent (each writes its own hC block and hlam
   entries), so a team of host threads shares them: at 3m = 36 one
   eigensolve takes about 2 ms, against a few ms of GPU work per half-batch
   that it has to hide behind. */
static int ritz_host_lobpcg(mx_solver *s, int j0, int j1, const int *plist)
{
  int status = MX_OK;
#pragma omp parallel for schedule(dynamic, 1) if (j1 - j0 > 1)
  for (int j = j0; j < j1; j++) {
    const int e = lobpcg_problem(s, j, plist ? plist[j] : j);
    if (e) {
#pragma omp critical(mx_ritz_status)
      if (!status) status = e;
    }
  }
  return status;
}

/* Launches the defects (K + (1 - lam) M) e of nv listed vectors (dvlist:
   device, null = 0..nv-1) into slots slot0 .. slot0 + nv - 1 of r (next
   right-hand sides, MG precision) and their norms to dnorm, without waiting. */
static int defect_launch(mx_solver *s, int nv, const int *dvlist, double *E, void *r, int slot0)
{
  const int L = s->L;
  const long long n = edim(s, L);
  mxh_bloch bl = batch_bloch(s);
  bl.vmap = dvlist;
  mxh_vecs ev = V(E, n);
  ev.map = dvlist;
  char *rs = (char *)r + (size_t)slot0 * n * (s->prec == MXH_F64 ? 16 : 8);
  CK(mxh_defect(s->prec, &s->lv[L], &bl, nv, ev, V(rs, n), s->eta64[L], s->dlam, s->dnorm + 2 * (size_t)slot0,
                s->dwork));
  s->st.launches += 2;
  return MX_OK;
}

/* Rayleigh-Ritz over the blocks of the listed problems (plist on the host,
   dplist on the device; null: problems 0..na-1): Gram sums, the small
   eigenproblems on the host, Eout = Ein C; then the defects of Eout's
   vectors (the m of each listed problem in list order: dvlist, nvl = na m)
   to r and their norms to hnorm.  With device lists and two or more
   problems the list is done in two halves, so that the host's eigensolves
   of one half run while the GPU sums the other half's Gram products or
   updates and checks the first half. */
static int ritz_defect(mx_solver *s, int na, const int *plist, const int *dplist, int nvl, const int *dvlist,
                       double *Ein, double *Eout, void *r)
{
  const int m = s->cfg.block, L = s->L, mg = s->vpp, np = mg * (mg + 1) / 2;
  const size_t csz = s->lobpcg ? (size_t)mg * 2 * m : (size_t)m * m; /* coefficients per problem */
  const long long n = edim(s, L);
  mxh_bloch bl = batch_bloch(s);
  const int split = na >= 2 && dplist && dvlist;
  const int nh[2] = {split ? (na + 1) / 2 : na, split ? na / 2 : 0};
  if (s->lobpcg) {
    /* the Gram's columns per problem: E, the W of the vectors preconditioned
       this step (a locked vector's W slot is not written) and P once it holds
       a direction; the Ritz step never uses the products it leaves out */
    for (int j = 0; j < na; j++) {
      const int p = plist ? plist[j] : j;
      int *cl = s->hcols + (size_t)p * (mg + 1), c = 0;
      for (int i = 0; i < m; i++) cl[1 + c++] = i;
      for (int i = 0; i < m; i++)
        if (s->hwvalid[p * m + i]) cl[1 + c++] = m + i;
      if (s->pvalid)
        for (int i = 0; i < m; i++) cl[1 + c++] = 2 * m + i;
      cl[0] = c;
    }
    CK(mxh_h2d(s->dcols, s->hcols, sizeof(int) * (size_t)(mg + 1) * s->last_nprob));
  }
  for (int h = 0, j0 = 0; h < 2 && nh[h]; j0 += nh[h], h++) {
    const int *pm = dplist ? dplist + j0 : NULL;
    double *gout = s->dgram + 4 * (size_t)np * j0;
    CK(s->lobpcg ? mxh_gram_cols(&s->lv[L], &bl, nh[h], pm, s->dcols, mg, V(Ein, n), s->eta64[L], gout, s->gramwork)
                 : mxh_gram(&s->lv[L], &bl, nh[h], pm, mg, V(Ein, n), s->eta64[L], gout, s->gramwork));
    CK(mxh_d2h_async(s->hgram + 4 * (size_t)np * j0, s->dgram + 4 * (size_t)np * j0,
                     sizeof(double) * 4 * (size_t)np * nh[h]));
    CK(mxh_event_record(s->ev_gram[h]));
    s->st.launches += 3;
  }
  for (int h = 0, j0 = 0; h < 2 && nh[h]; j0 += nh[h], h++) {
    CK(mxh_event_sync(s->ev_gram[h]));
    int e = s->lobpcg ? ritz_host_lobpcg(s, j0, j0 + nh[h], plist) : ritz_host(s, j0, j0 + nh[h], plist);
    if (e) {
      mxh_sync();
      return e;
    }
    CK(mxh_h2d(s->dC + 2 * (size_t)j0 * csz, s->hC + (size_t)j0 * csz, sizeof(double complex) * csz * nh[h]));
    CK(mxh_h2d(s->dlam, s->hlam, sizeof(double) * (size_t)mg * s->last_nprob));
    if (s->lobpcg && !(m & 1)) /* [E W P] -> E_new (slots 0..m-1), P_new (slots 2m..3m-1) */
      CK(mxh_lobpcg_combine(n, nh[h], dplist ? dplist + j0 : NULL, m, V(Ein, n), V(Eout, n),
                            s->dC + 2 * (size_t)j0 * csz));
    else if (s->lobpcg)
      CK(mxh_combine_x(n, nh[h], dplist ? dplist + j0 : NULL, mg, 2 * m, m, m, V(Ein, n), mg, V(Eout, n), mg,
                       s->dC + 2 * (size_t)j0 * csz));
    else if (m > 16)  /* every input read once */
      CK(mxh_combine_x(n, nh[h], dplist ? dplist + j0 : NULL, m, m, m, 0, V(Ein, n), m, V(Eout, n), m,
                       s->dC + 2 * (size_t)j0 * csz));
    else
      CK(mxh_block_combine(n, nh[h], dplist ? dplist + j0 : NULL, m, V(Ein, n), V(Eout, n),
                           s->dC + 2 * (size_t)j0 * csz));
    s->st.launches++;
    if ((e = defect_launch(s, nh[h] * m, dvlist ? dvlist + j0 * m : NULL, Eout, r, j0 * m))) return e;
  }
  CK(mxh_d2h(s->hnorm, s->dnorm, sizeof(double) * 2 * nvl));
  return MX_OK;
}

/* The preconditioner of one block step for the ng vectors of the multigrid
   set (gmap: their batch-wide indices, bmap: their defect slots in rdef):
   b = gathered defects; x = V^{prec_steps}(b) from zero; e <- u - G(...) with
   u = e - x, the gradient lifting in place in E.  Host work here is launch
   set-up only. */
static int precondition_to(mx_solver *s, int ng, const mxh_bloch *blg, double *E, const int *dbmap,
                           const int *dgmap, double *W, const int *dwmap)
{
  const int L = s->L;
  const long long n = edim(s, L);
  int e;
  /* the finest level's row-marching smoother reads the defects where they
     were formed, through the set's slot map; otherwise they are gathered */
  mxh_vecs src = V(s->rdef, n), bfine = V(NULL, 0);
  src.map = dbmap;
  if (!s->cfg.phase_kernels && s->cfg.smoother_steps > 0 && march_rows_of(s, 0, L, ng) > 0) {
    bfine = src;
  } else {
    CK(mxh_gather(s->prec, n, ng, V(s->eb[L], n), src));
    s->st.launches++;
  }
  for (int k = 0; k < s->cfg.prec_steps; k++)
    if ((e = vcycle_edge(s, ng, blg, k == 0, bfine))) return e;
  return lift_to(s, ng, blg, E, s->ex[L], dgmap, W, dwmap);
}

static int precondition(mx_solver *s, int ng, const mxh_bloch *blg, double *E, const int *dbmap,
                        const int *dgmap)
{
  return precondition_to(s, ng, blg, E, dbmap, dgmap, NULL, NULL);
}

static void buffer_assignment(const mx_solver *s, void *p[4][MX_MAXL])
{
  for (int l = 0; l < MX_MAXL; l++) {
    p[0][l] = l <= s->L ? s->ex[l] : NULL;
    p[1][l] = l <= s->L ? s->ex2[l] : NULL;
    p[2][l] = l <= s->L ? s->nxv[l] : NULL;
    p[3][l] = l <= s->L ? s->nxv2[l] : NULL;
  }
}

/* precondition() through a HIP graph.  The region's kernels and arguments
   depend on the host state only through ng, which of dE and dU holds the
   block (E), the fixed list and block buffers, and which of the two x
   buffers of each level is current; the
   smoothers swap those per launch, an even number of times per V-cycle (the
   same sweep sequence down and up), so the assignment after the region
   equals the one before and a replay leaves the host state as valid as an
   eager run.  A changed ng or assignment is captured again (the graph is
   updated in place when only launch shapes and arguments differ). */
static int precondition_graph(mx_solver *s, int ng, const mxh_bloch *blg, double *E,
                              const int *dbmap, const int *dgmap)
{
  void *cur[4][MX_MAXL];
  buffer_assignment(s, cur);
  if (!s->graph || s->graph_ng != ng || s->graph_E != E || memcmp(cur, s->graph_ptr, sizeof(cur))) {
    const mx_stats before = s->st;
    CK(mxh_capture_begin());
    int e = precondition(s, ng, blg, E, dbmap, dgmap);
    if (e) {
      mxh_capture_abort();
      return e;
    }
    void *after[4][MX_MAXL];
    buffer_assignment(s, after);
    if (memcmp(cur, after, sizeof(cur))) {
      mxh_capture_abort();
      return MX_EDEVICE;
    }
    if (mxh_capture_end(&s->graph)) {
      s->graph = NULL;
      return MX_EDEVICE;
    }
    s->graph_launches = s->st.launches - before.launches;
    s->graph_edge = s->st.edge_vcycles - before.edge_vcycles;
    s->graph_node = s->st.node_vcycles - before.node_vcycles;
    s->st = before; /* capturing ran nothing */
    s->st.graph_captures++;
    s->graph_ng = ng;
    s->graph_E = E;
    memcpy(s->graph_ptr, cur, sizeof(cur));
  }
  CK(mxh_graph_launch(s->graph));
  s->st.launches += s->graph_launches;
  s->st.edge_vcycles += s->graph_edge;
  s->st.node_vcycles += s->graph_node;
  s->st.graph_replays++;
  return MX_OK;
}

static double now_seconds(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

static int setup_problems(mx_solver *s, int nprob, const double *kx, const double *ky)
{
  const int P = nprob, ne0 = s->ne0, nn0 = s->nn0, n0 = s->cfg.n0;
  for (int p = 0; p < P; p++) {
    double complex xf = cos(kx[p]) + I * sin(kx[p]);
    double complex yf = cos(ky[p]) + I * sin(ky[p]);
    s->hbloch[4 * p + 0] = creal(xf);
    s->hbloch[4 * p + 1] = cimag(xf);
    s->hbloch[4 * p + 2] = creal(yf);
    s->hbloch[4 * p + 3] = cimag(yf);
    s->gamma[p] = (xf == 1.0 && yf == 1.0);
  }
  CK(mxh_h2d(s->dbloch, s->hbloch, sizeof(double) * 4 * P));
  /* per-k coarse rebuild (exp_bloch2d.c:799-809): assembled and factored on
     the device, one workgroup per problem, for coarse grids whose edge matrix
     fits in LDS (2 n0^2 <= 64, the reference's n0 = 4 included); larger
     coarse grids are assembled and factored here and uploaded */
  if (ne0 <= MXH_COARSE_SETUP_MAX) {
    CK(mxh_coarse_setup(n0, P, s->dbloch, s->eta64[0], s->dAe, s->dAn, s->dinfo));
    CK(mxh_d2h(s->hinfo, s->dinfo, sizeof(int) * P));
    for (int p = 0; p < P; p++)
      if (s->hinfo[p]) return MX_ELAPACK;
    return MX_OK;
  }
  for (int p = 0; p < P; p++) {
    const double complex xf = s->hbloch[4 * p] + I * s->hbloch[4 * p + 1];
    const double complex yf = s->hbloch[4 * p + 2] + I * s->hbloch[4 * p + 3];
    double complex *Ae = s->hAe + (size_t)ne0 * ne0 * p;
    double complex *An = s->hAn + (size_t)nn0 * nn0 * p;
    mxd_dense_edge(n0, n0, 1.0 / n0, 1.0 / n0, xf, yf, s->hEtaL[0], 1.0, 1.0, Ae, ne0);
    if (mxd_potrf(ne0, Ae)) return MX_ELAPACK;
    mxd_dense_node(n0, n0, 1.0 / n0, 1.0 / n0, xf, yf, 1, An, nn0);
    if (mxd_potrf(nn0, An)) return MX_ELAPACK;
  }
  CK(mxh_h2d(s->dAe, s->hAe, sizeof(double complex) * (size_t)ne0 * ne0 * P));
  CK(mxh_h2d(s->dAn, s->hAn, sizeof(double complex) * (size_t)nn0 * nn0 * P));
  return MX_OK;
}

int mx_solve(mx_solver *s, const mx_batch *b, mx_result *r)
{
  if (!s || !b || !r || b->nprob < 1 || b->nprob > s->cfg.max_problems || b->neig < 1 ||
      b->neig > s->cfg.block || !b->kx || !b->ky || !r->lambda || !r->relres || !r->absres || !r->iters ||
      !r->converged || b->maxit < 0 || !(b->tol >= 0.0))
    return MX_EINVAL;
  for (int p = 0; p < b->nprob; p++)
    if (!isfinite(b->kx[p]) || !isfinite(b->ky[p])) return MX_EINVAL;
  const double t0 = now_seconds();
  const int P = b->nprob, m = s->cfg.block, nv = P * m, L = s->L;
  const long long n = edim(s, L);
  int e;
  s->last_nprob = P;
  if ((e = setup_problems(s, P, b->kx, b->ky))) return e;
  mxh_bloch bl = batch_bloch(s);
  if (s->cfg.autotune && !s->tuned && (e = autotune(s, nv, &bl))) return e;

  const int vpp = s->vpp;
  if (s->lobpcg && (m & 1)) {
    /* odd blocks: the general block update reads every input of [E W P]
       (times a zero coefficient for the columns the Ritz step dropped), so
       the search directions W and P start as zeros in both buffers; the
       even-block update and the Gram read only the columns in use (E, the W
       of unlocked vectors, P once it holds a direction) */
    CK(mxh_memset(s->dE, 0, 16 * (size_t)n * vpp * P));
    CK(mxh_memset(s->dU, 0, 16 * (size_t)n * vpp * P));
  }
  /* start block in the E slots (problem p's vectors p vpp .. p vpp + m - 1;
     the random stream is indexed by p m + i in both layouts) */
  if (b->E0) {
    if (vpp == m)
      CK(mxh_h2d(s->dE, b->E0, 16 * (size_t)n * nv));
    else
      for (int p = 0; p < P; p++)
        CK(mxh_h2d(s->dE + 2 * (size_t)p * vpp * n, b->E0 + 2 * (size_t)p * m * n, 16 * (size_t)n * m));
  } else if (vpp == m) {
    CK(mxh_random(MXH_F64, n, nv, V(s->dE, n), b->seed, 0));
    s->st.launches++;
  } else {
    for (int p = 0; p < P; p++) CK(mxh_random(MXH_F64, n, m, V(s->dE + 2 * (size_t)p * vpp * n, n), b->seed, p * m));
    s->st.launches += P;
  }
  /* host lists: active problems, vectors of active problems (defect slots),
     multigrid set (slot in the defect output, block-buffer vector of e and,
     with LOBPCG, of its search direction w); block-buffer index of problem
     p's vector i: p vpp + i */
  int *plist = s->hlist, *vlist = plist + P, *bmap = vlist + nv, *gmap = bmap + nv, *wmap = gmap + nv;
  int *dplist = s->dlist, *dvlist = dplist + P, *dbmap = dvlist + nv, *dgmap = dbmap + nv, *dwmap = dgmap + nv;
  int *slot = s->hslot;
  for (int p = 0; p < P; p++) plist[p] = p;
  for (int p = 0, v = 0; p < P; p++)
    for (int i = 0; i < m; i++, v++) {
      vlist[v] = gmap[v] = p * vpp + i;
      slot[p * vpp + i] = v;
    }
  for (int v = 0; v < nv; v++) s->hwvalid[v] = 0;
  s->pvalid = 0;
  int na = P, nvl = nv;

  CK(mxh_h2d(dplist, plist, sizeof(int) * (size_t)(gmap + nv - plist)));
  {
    mxh_bloch bls = bl;
    bls.vmap = dgmap;
    if ((e = center_gamma(s, P, NULL, s->dE))) return e;
    if ((e = lift(s, nv, &bls, s->dE, NULL, dgmap))) return e;
    if ((e = center_gamma(s, P, NULL, s->dE))) return e;
  }
  if ((e = ritz_defect(s, P, plist, dplist, nv, dvlist, s->dE, s->dU, s->rdef))) return e;
  double *E = s->dU, *U = s->dE;

  const size_t nslots = (size_t)vpp * P;
  double *rel = (double *)malloc(sizeof(double) * nslots), *ab = (double *)malloc(sizeof(double) * nslots);
  int *done = (int *)calloc(P, sizeof(int));
  /* the block buffer that holds each problem's final Ritz vectors: a problem
     leaves the active set at a step of either parity, and later steps only
     write the active problems' vectors into the alternating buffers */
  double **held = (double **)malloc(sizeof(double *) * P);
  if (!rel || !ab || !done || !held) { free(rel); free(ab); free(done); free(held); return MX_ENOMEM; }
  for (int p = 0; p < P; p++) { r->converged[p] = 0; r->iters[p] = b->maxit; }
  int status = MX_OK;
  for (int it = 0;; it++) {
    /* residuals of the vectors of the active problems (defect slots) */
    for (int q = 0; q < nvl; q++) {
      const int v = vlist[q];
      double rn = sqrt(s->hnorm[2 * q]), mn = sqrt(s->hnorm[2 * q + 1]);
      ab[v] = rn;
      rel[v] = rn / (fabs(s->hlam[v] - 1.0) * mn);
      slot[v] = q;
    }
    int nna = 0;
    for (int j = 0; j < na && !status; j++) {
      const int p = plist[j];
      double worst = 0.0;
      for (int i = 0; i < b->neig; i++) {
        double q = rel[p * vpp + i];
        if (!isfinite(q)) status = MX_ENOTFINITE;
        else if (q > worst) worst = q;
      }
      if (worst <= b->tol || it >= b->maxit || status) {
        done[p] = 1;
        held[p] = E;
        r->converged[p] = !status && worst <= b->tol;
        r->iters[p] = it;
        for (int i = 0; i < m; i++) {
          r->lambda[p * m + i] = s->hlam[p * vpp + i] - 1.0;
          r->relres[p * m + i] = rel[p * vpp + i];
          r->absres[p * m + i] = ab[p * vpp + i];
        }
      } else {
        plist[nna++] = p;
      }
    }
    if (status || nna == 0) break;
    na = nna;

    /* multigrid set: unconverged wanted pairs and the extra block vectors;
       converged pairs are locked (kept in the Ritz basis, not preconditioned) */
    int ng = 0;
    for (int j = 0; j < na; j++)
      for (int i = 0; i < m; i++) {
        const int p = plist[j], g = p * vpp + i;
        const int pre = i >= b->neig || rel[g] > b->tol || !s->cfg.lock;
        s->hwvalid[p * m + i] = pre;
        if (pre) {
          gmap[ng] = g;
          wmap[ng] = g + m;
          bmap[ng] = slot[g];
          ng++;
        }
      }
    nvl = 0;
    for (int j = 0; j < na; j++)
      for (int i = 0; i < m; i++) vlist[nvl++] = plist[j] * vpp + i;
    /* the lists sit back to back (P + 4 nv ints): one upload up to the end
       of the last list's used part */
    CK(mxh_h2d(dplist, plist, sizeof(int) * (size_t)((s->lobpcg ? wmap : gmap) + ng - plist)));
    s->st.active_vectors += ng;

    /* compacted right-hand sides of the multigrid set, x = V^{prec_steps}(b)
       from zero; PINVIT: u = e - x, gradient lifting back into E (locked
       vectors stay in E as they are); LOBPCG: the lifted u - e into the W
       slots */
    mxh_bloch blg = bl;
    blg.vmap = dgmap;
    if (s->lobpcg) {
      e = precondition_to(s, ng, &blg, E, dbmap, dgmap, E, dwmap);
      if (e) { status = e; break; }
      /* Gamma point: the lifted update's constant fields (W slots) */
      if ((e = center_gamma(s, na, plist, E + 2 * (size_t)m * n))) { status = e; break; }
    } else {
      const int use_graph = s->cfg.graphs && !s->cfg.phase_kernels && s->cfg.smoother_steps > 0;
      e = use_graph ? precondition_graph(s, ng, &blg, E, dbmap, dgmap) : precondition(s, ng, &blg, E, dbmap, dgmap);
      if (e) { status = e; break; }
      if ((e = center_gamma(s, na, plist, E))) { status = e; break; }
    }
    /* Ritz on E -> U, which becomes the block; defects of the active
       problems' vectors */
    if ((e = ritz_defect(s, na, plist, dplist, nvl, dvlist, E, U, s->rdef))) { status = e; break; }
    /* from here on the P slots hold [W P] C_WP with this step's W */
    s->pvalid = 1;
    double *t = E;
    E = U;
    U = t;
    s->st.block_steps += na;
  }
  /* keep every problem's Ritz vectors in dE for mx_get_vectors /
     mx_eta_gradient: each problem's m vectors are contiguous, copied from the
     buffer that held the block when the problem finished (problems still
     active after an error: the current block) */
  int cerr = 0;
  for (int p = 0; p < P && !cerr; p++) {
    const double *src = done[p] ? held[p] : E;
    if (src != s->dE)
      cerr = mxh_d2d(s->dE + 2 * (size_t)p * vpp * n, src + 2 * (size_t)p * vpp * n, 16 * (size_t)n * m);
  }
  free(rel);
  free(ab);
  free(done);
  free(held);
  CK(cerr);
  CK(mxh_sync());
  s->st.solve_seconds = now_seconds() - t0;
  return status;
}

int mx_get_vectors(mx_solver *s, int p, double *E)
{
  if (!s || !E || p < 0 || p >= s->last_nprob) return MX_EINVAL;
  const long long n = edim(s, s->L);
  const int m = s->cfg.block;
  CK(mxh_d2h(E, s->dE + 2 * (size_t)p * s->vpp * n, 16 * (size_t)m * n));
  return MX_OK;
}

int mx_eta_gradient(mx_solver *s, int p, int first, int count, double *grad)
{
  const int m = s ? s->cfg.block : 0;
  if (!s || !grad || p < 0 || p >= s->last_nprob || first < 0 || count < 1 || first + count > m)
    return MX_EINVAL;
  const int L = s->L;
  const long long n = edim(s, L), N = ndim(s, L);
  double w[48];
  double *Ep = s->dE + 2 * ((size_t)p * s->vpp + first) * n;
  mxh_bloch bl;
  memset(&bl, 0, sizeof(bl));
  bl.xr = s->hbloch[4 * p + 0];
  bl.xi = s->hbloch[4 * p + 1];
  bl.yr = s->hbloch[4 * p + 2];
  bl.yi = s->hbloch[4 * p + 3];
  /* 1 / (e^H M e) from the Gram kernel on the cluster */
  const int np = count * (count + 1) / 2;
  CK(mxh_gram(&s->lv[L], &bl, 1, NULL, count, V(Ep, n), s->eta64[L], s->dgram, s->gramwork));
  double *hg = (double *)malloc(sizeof(double) * 4 * np);
  if (!hg) return MX_ENOMEM;
  if (mxh_d2h(hg, s->dgram, sizeof(double) * 4 * np)) { free(hg); return MX_EDEVICE; }
  for (int i = 0; i < count; i++) w[i] = 1.0 / hg[4 * (i * (i + 1) / 2 + i) + 2];
  free(hg);
  double *dw = s->dwork;
  double *dg = (double *)mxh_malloc(sizeof(double) * N);
  if (!dg) return MX_ENOMEM;
  int e = mxh_h2d(dw, w, sizeof(double) * count);
  if (!e) e = mxh_memset(dg, 0, sizeof(double) * N);
  if (!e) e = mxh_curl2(&s->lv[L], &bl, count, V(Ep, n), dw, dg);
  if (!e) e = mxh_d2h(grad, dg, sizeof(double) * N);
  mxh_free(dg);
  return e ? MX_EDEVICE : MX_OK;
}

int mx_get_tuning(const mx_solver *s, int *edge_rows, int *node_rows)
{
  if (!s || !edge_rows || !node_rows) return MX_EINVAL;
  for (int l = 0; l <= s->L; l++) {
    edge_rows[l] = l == 0 ? 0 : march_rows_of(s, 0, l, s->nvmax);
    node_rows[l] = l == 0 ? 0 : march_rows_of(s, 1, l, s->nvmax);
  }
  return MX_OK;
}

/* Rayleigh-Ritz extrapolation of start blocks (rayleigh_extrapolation,
   exp_blochsurf.c:332-406; the '1' and '2' cases of exp_bloch2d.c:825-894),
   per problem on the device: the sources * ieigs basis vectors are
   M-orthonormalised by Householder reflections (mxh_orthonormalize, the
   reference's orthonormalize_edge2d), the Ritz matrix of K + M is formed over
   the orthonormal basis (the Gram kernel's pairs, the reference's
   energyprod_edge2d(1, 1, .)), its Hermitian eigenproblem is solved on the
   host (the reference's syev_matrix) and the lowest `block` Ritz vectors are
   formed on the device (mxh_combine_x, the reference's add_edge2d loop). */
int mx_extrapolate(mx_solver *s, const mx_batch *b, int sources, int ieigs, const double *H, double *E0,
                   double *lambda)
{
  if (!s || !b || !H || !E0 || sources < 1 || ieigs < 1) return MX_EINVAL;
  const int m = s->cfg.block, dim = sources * ieigs, L = s->L;
  /* the Householder units sit on x-edges of odd rows: nx (ny / 2) of them */
  if (b->nprob < 1 || b->nprob > s->cfg.max_problems || !b->kx || !b->ky || dim < m || dim > 48 ||
      (long long)dim > (long long)s->lv[L].nx * (s->lv[L].ny / 2))
    return MX_EINVAL;
  const long long n = edim(s, L);
  const size_t vb = 16 * (size_t)n; /* bytes per complex FP64 vector */
  const int np = dim * (dim + 1) / 2;
  const double sk = 1.0 / (s->lv[L].hx * s->lv[L].hy);
  double *X = (double *)mxh_malloc(vb * dim), *Vv = (double *)mxh_malloc(vb * dim);
  double *tmp = (double *)mxh_malloc(vb), *out = (double *)mxh_malloc(vb * m);
  double *tau = (double *)mxh_malloc(16 * (size_t)dim), *work = (double *)mxh_malloc(mxh_orthonormalize_work_bytes(dim));
  double *dg = (double *)mxh_malloc(sizeof(double) * 4 * np), *gw = (double *)mxh_malloc(mxh_gram_work_bytes(dim, 1));
  double *dC = (double *)mxh_malloc(16 * (size_t)dim * m);
  double *hg = (double *)malloc(sizeof(double) * 4 * np);
  double complex *A = (double complex *)malloc(sizeof(double complex) * dim * dim);
  double *w = (double *)malloc(sizeof(double) * dim);
  int err = (!X || !Vv || !tmp || !out || !tau || !work || !dg || !gw || !dC || !hg || !A || !w) ? MX_ENOMEM : MX_OK;
  for (int p = 0; p < b->nprob && !err; p++) {
    const double xr = cos(b->kx[p]), xi = sin(b->kx[p]), yr = cos(b->ky[p]), yi = sin(b->ky[p]);
    mxh_bloch bl;
    memset(&bl, 0, sizeof(bl));
    bl.xr = xr;
    bl.xi = xi;
    bl.yr = yr;
    bl.yi = yi;
    if (mxh_h2d(X, H + 2 * (size_t)p * dim * n, vb * dim) ||
        mxh_orthonormalize(&s->lv[L], xr, xi, yr, yi, s->eta64[L], dim, X, Vv, tmp, tau, work) ||
        mxh_gram(&s->lv[L], &bl, 1, NULL, dim, V(Vv, n), s->eta64[L], dg, gw) ||
        mxh_d2h(hg, dg, sizeof(double) * 4 * np)) {
      err = MX_EDEVICE;
      break;
    }
    for (int q = 0; q < 4 * np; q++)
      if (!isfinite(hg[q])) err = MX_ENOTFINITE;
    if (err) break;
    /* lower triangle of the Ritz matrix (K + M on the orthonormal basis) */
    for (int a = 0; a < dim; a++)
      for (int c = 0; c <= a; c++) {
        const double *g = hg + 4 * (a * (a + 1) / 2 + c);
        A[a + (size_t)c * dim] = sk * (g[0] + I * g[1]) + (g[2] + I * g[3]);
        A[c + (size_t)a * dim] = conj(A[a + (size_t)c * dim]);
      }
    if (mxd_heev(dim, A, w)) {
      err = MX_ELAPACK;
      break;
    }
    if (lambda)
      for (int j = 0; j < m; j++) lambda[(size_t)p * m + j] = w[j] - 1.0;
    if (mxh_h2d(dC, A, 16 * (size_t)dim * m) ||
        mxh_combine_x(n, 1, NULL, dim, m, m, 0, V(Vv, n), dim, V(out, n), m, dC) ||
        mxh_d2h(E0 + 2 * (size_t)p * m * n, out, vb * m))
      err = MX_EDEVICE;
  }
  mxh_free(X);
  mxh_free(Vv);
  mxh_free(tmp);
  mxh_free(out);
  mxh_free(tau);
  mxh_free(work);
  mxh_free(dg);
  mxh_free(gw);
  mxh_free(dC);
  free(hg);
  free(A);
  free(w);
  return err;
}

int mx_get_stats(const mx_solver *s, mx_stats *st)
{
  if (!s || !st) return MX_EINVAL;
  *st = s->st;
  return MX_OK;
}

int mx_time_kernel(mx_solver *s, int kind, int reps, double *avg_ms, double *bytes)
{
  if (!s || reps < 1 || !avg_ms || !bytes || s->last_nprob < 1) return MX_EINVAL;
  const int L = s->L, nv = s->last_nprob * s->cfg.block;
  mxh_bloch bl = batch_bloch(s);
  void *e0, *e1;
  CK(mxh_event_create(&e0));
  if (mxh_event_create(&e1)) {
    mxh_event_destroy(e0);
    return MX_EDEVICE;
  }
  const double E = (double)edim(s, L), C = (double)ndim(s, L);
  const double c = (double)s->csz, r = s->prec == MXH_F64 ? 8.0 : 4.0;
  double by = 0.0;
  int err = 0;
  err |= mxh_sync();
  err |= mxh_event_record(e0);
  for (int k = 0; k < reps && !err; k++) {
    switch (kind) {
    case 0:
      err |= mxh_afw_phase(s->prec, &s->lv[L], &bl, nv, V(s->eb[L], edim(s, L)), V(s->ex[L], edim(s, L)),
                           s->etaP[L], 1.0, 1.0, k & 3);
      break;
    case 1:
      err |= vcycle_edge(s, nv, &bl, 0, V(NULL, 0));
      break;
    case 2:
      err |= vcycle_node(s, nv, &bl, 0);
      break;
    case 3:
      err |= mxh_node_gs_phase(s->prec, &s->lv[L], &bl, nv, V(s->nb[L], ndim(s, L)), V(s->nxv[L], ndim(s, L)),
                               1.0, k & 3);
      break;
    case 4:
      err |= mxh_afw_tile(s->prec, &s->lv[L], &bl, nv, V(s->eb[L], edim(s, L)), V(s->ex[L], edim(s, L)),
                          V(s->ex2[L], edim(s, L)), s->etaP[L], 1, 0, V(NULL, 0), V(NULL, 0), V(NULL, 0));
      swap_ptr(&s->ex[L], &s->ex2[L]);
      break;
    case 5:
      err |= mxh_node_tile(s->prec, &s->lv[L], &bl, nv, V(s->nb[L], ndim(s, L)), V(s->nxv[L], ndim(s, L)),
                           V(s->nxv2[L], ndim(s, L)), 1, 0, V(NULL, 0), V(NULL, 0), V(NULL, 0));
      swap_ptr(&s->nxv[L], &s->nxv2[L]);
      break;
    case 6:
    case 10:
      err |= mxh_afw_march(s->prec, &s->lv[L], &bl, nv, V(s->eb[L], edim(s, L)), V(s->ex[L], edim(s, L)),
                           V(s->ex2[L], edim(s, L)), s->etaP[L], 0, V(NULL, 0), V(NULL, 0), V(NULL, 0),
                           s->cfg.march_rows, kind == 10 ? 2 : 1);
      swap_ptr(&s->ex[L], &s->ex2[L]);
      break;
    case 7:
    case 11:
      err |= mxh_node_march(s->prec, &s->lv[L], &bl, nv, V(s->nb[L], ndim(s, L)), V(s->nxv[L], ndim(s, L)),
                            V(s->nxv2[L], ndim(s, L)), 0, V(NULL, 0), V(NULL, 0), V(NULL, 0), s->cfg.march_rows,
                            kind == 11 ? 2 : 1);
      swap_ptr(&s->nxv[L], &s->nxv2[L]);
      break;
    case 8:
      err |= mxh_d2d(s->ex2[L], s->ex[L], s->csz * (size_t)edim(s, L) * nv);
      break;
    case 9:
      err |= mxh_gram(&s->lv[L], &bl, s->last_nprob, NULL, s->vpp, V(s->dE, edim(s, L)), s->eta64[L],
                      s->dgram, s->gramwork);
      break;
    case 12:
    case 13:
    case 14: {
      /* the finest level's two-sweep passes of the solve's V(2,2) cycles,
         with the launch shape the solve uses for the whole batch */
      static const int pass_mode[3] = {10, 8, 5};
      const int mode = pass_mode[kind - 12], rows = march_rows_of(s, 0, L, nv);
      const mxh_vecs none = V(NULL, 0), xcv = V(s->ex[L - 1], edim(s, L - 1));
      if (rows < 1 || L < 1 || s->cfg.smoother_steps != 2 || !s->cfg.march_fuse) {
        err = 1;
        break;
      }
      err |= mxh_afw_march(s->prec, &s->lv[L], &bl, nv, V(s->eb[L], edim(s, L)), V(s->ex[L], edim(s, L)),
                           V(s->ex2[L], edim(s, L)), s->etaP[L], mode, mode == 5 ? xcv : none,
                           mode == 5 ? none : V(s->eb[L - 1], edim(s, L - 1)), mode == 5 ? none : xcv, rows, 2);
      swap_ptr(&s->ex[L], &s->ex2[L]);
      break;
    }
    case 15:
    case 16:
    case 17: {
      /* the nodal counterparts of kinds 12-14 (the lifting's node V-cycles) */
      static const int pass_mode[3] = {10, 8, 5};
      const int mode = pass_mode[kind - 15], rows = march_rows_of(s, 1, L, nv);
      const mxh_vecs none = V(NULL, 0), xcv = V(s->nxv[L - 1], ndim(s, L - 1));
      if (rows < 1 || L < 1 || s->cfg.smoother_steps != 2 || !s->cfg.march_fuse) {
        err = 1;
        break;
      }
      err |= mxh_node_march(s->prec, &s->lv[L], &bl, nv, V(s->nb[L], ndim(s, L)), V(s->nxv[L], ndim(s, L)),
                            V(s->nxv2[L], ndim(s, L)), mode, mode == 5 ? xcv : none,
                            mode == 5 ? none : V(s->nb[L - 1], ndim(s, L - 1)), mode == 5 ? none : xcv, rows, 2);
      swap_ptr(&s->nxv[L], &s->nxv2[L]);
      break;
    }
    default:
      err = 1;
    }
  }
  err |= mxh_event_record(e1);
  float ms = 0.0f;
  err |= mxh_event_elapsed_ms(e0, e1, &ms);
  mxh_event_destroy(e0);
  mxh_event_destroy(e1);
  if (err) return MX_EDEVICE;
  *avg_ms = ms / reps;
  /* algorithmic bytes: each stream counted once */
  switch (kind) {
  case 0: /* one colour phase: read b and x of the half of the DOFs it owns,
             read the neighbour x it needs (all x once), write its half, eta */
    by = nv * (E * c * 0.5 + E * c + E * c * 0.5) + C * r;
    break;
  case 1: { /* V(nu,nu), fused-ideal: per level 2 nu sweeps (3 E c + C r each),
               residual/restriction (2 E c + C r + 0.25 E c), prolongation
               (0.25 E c + 2 E c) */
    double tot = 0.0;
    for (int l = 1; l <= L; l++) {
      double El = (double)edim(s, l), Cl = (double)ndim(s, l);
      tot += nv * (2 * s->cfg.smoother_steps * 3 * El * c + 2.25 * El * c + 2.25 * El * c) +
             (2 * s->cfg.smoother_steps + 1) * Cl * r;
    }
    by = tot;
    break;
  }
  case 2: {
    double tot = 0.0;
    for (int l = 1; l <= L; l++) {
      double Nl = (double)ndim(s, l);
      tot += nv * (2 * s->cfg.smoother_steps * 3 * Nl * c + 2.25 * Nl * c + 2.25 * Nl * c);
    }
    by = tot;
    break;
  }
  case 3: /* node colour phase: read b of its quarter, all x once, write quarter */
    by = nv * (C * c * 0.25 + C * c + C * c * 0.25);
    break;
  case 4: /* one fused AFW sweep: read x, b, eta once, write x once */
    by = nv * 3 * E * c + C * r;
    break;
  case 5: /* one fused nodal sweep: read x, b, write x */
  case 7:
    by = nv * 3 * C * c;
    break;
  case 6: /* as kind 4 */
    by = nv * 3 * E * c + C * r;
    break;
  case 10: /* two sweeps, each counted as kind 4 */
    by = 2 * (nv * 3 * E * c + C * r);
    break;
  case 11:
    by = 2 * nv * 3 * C * c;
    break;
  case 8: /* copy: read + write */
    by = nv * 2 * E * c;
    break;
  case 9: /* Gram of every problem's block (LOBPCG: [E W P]): each FP64 vector read once, eta */
    by = (double)s->last_nprob * s->vpp * E * 16.0 + C * 8.0;
    break;
  /* the solve's fused passes, each part counted as its own stream: a sweep
     3 E c + C r (the first sweep of a zero start reads no x: 2 E c + C r),
     the residual 3 E c + C r, the restriction E c + E c / 4, the
     prolongation E c / 4 + 2 E c */
  case 12: /* zero start, two sweeps, residual, restriction */
    by = nv * (2 * E * c + 3 * E * c + 3 * E * c + 1.25 * E * c) + 3 * C * r;
    break;
  case 13: /* two sweeps, residual, restriction */
    by = nv * (3 * E * c + 3 * E * c + 3 * E * c + 1.25 * E * c) + 3 * C * r;
    break;
  case 14: /* prolongation, two sweeps */
    by = nv * (2.25 * E * c + 3 * E * c + 3 * E * c) + 2 * C * r;
    break;
  /* nodal passes, each part its own stream: a sweep 3 N c (a zero start
     reads no x: 2 N c), the residual 3 N c, the restriction 1.25 N c, the
     prolongation 2.25 N c, N = nodes = C */
  case 15:
    by = nv * (2 * C * c + 3 * C * c + 3 * C * c + 1.25 * C * c);
    break;
  case 16:
    by = nv * (3 * C * c + 3 * C * c + 3 * C * c + 1.25 * C * c);
    break;
  case 17:
    by = nv * (2.25 * C * c + 3 * C * c + 3 * C * c);
    break;
  }
  *bytes = by;
  return MX_OK;
}
