/*
 * maxwell_amd.h -- entry points added on top of the reference's module API.
 *
 * The reference runs the Bloch path one k-point and one vector at a time:
 * exp_bloch2d.c:956-1053 and exp_blochsurf.c:137-227 call pinvit_edge2d
 * (edge2d.c:6298-6328) per vector, then build the Ritz system from
 * energyprod/massprod pairs and solve it with zhegv.  These entry points do
 * the same mathematics for whole blocks and for many Bloch parameters per GPU
 * launch.  All return 0 on success or a nonzero status (MX_E*), never abort.
 *
 * Plain C types only: sizes are ints, vectors are host arrays of interleaved
 * (re, im) doubles in the dense ordering of densefrom_edge2d (x-edges first,
 * then y-edges, row-major), eigenvalues are those of K (the reference prints
 * lambda - 1 of its K + M pencil, exp_bloch2d.c:1227-1235).
 */
#ifndef MAXWELL_AMD_H
#define MAXWELL_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MX_OK = 0,
  MX_EINVAL = 1,   /* bad argument */
  MX_ENOMEM = 2,   /* device or host allocation failed */
  MX_EDEVICE = 3,  /* HIP error */
  MX_ELAPACK = 4,  /* small dense eigenproblem failed (Gram not definite) */
  MX_ENOTFINITE = 5 /* NaN or Inf in a residual */
};

enum { MX_PREC_F64 = 0, MX_PREC_MIXED = 1 };
enum { MX_METHOD_PINVIT = 0, MX_METHOD_LOBPCG = 1 };

typedef struct {
  int n0;             /* coarse grid intervals per side (maxwell_coarse) */
  int levels;         /* refinements L; finest grid n0 << L (maxwell_levels) */
  int precision;      /* MX_PREC_F64, or MX_PREC_MIXED: FP32 multigrid,
                         FP64 residuals, Gram matrices and Ritz step */
  int block;          /* vectors per problem, eigs + throwaway (<= 32) */
  int max_problems;   /* Bloch problems solved together per launch */
  int smoother_steps; /* nu of V(nu,nu) (reference: 2) */
  int prec_steps;     /* edge V-cycles per PINVIT step (exp_bloch2d.c:233: 2) */
  int gradient_steps; /* node V-cycles per lifting (exp_bloch2d.c:234: 5) */
  int phase_kernels;  /* 0: tile-fused smoother kernels (default); 1: one
                         launch per colour phase (same arithmetic, for checks) */
  int lock;           /* 1: converged wanted pairs are locked -- kept in the
                         Ritz basis but not preconditioned -- until their
                         residual exceeds tol again; 0: every vector is
                         updated every step (the reference's loop) */
  int march_min;      /* levels with n >= march_min smooth with the row-
                         marching kernels; so does the level of n/2 when a
                         launch has at least three marching workgroups per
                         compute unit; smaller ones use square tiles
                         (0: square tiles everywhere; 256 is fastest on
                         MI355X for 512^2 batches) */
  int march_rows;     /* owned rows per marching workgroup (0: 128; 192 is
                         fastest on MI355X for 512^2 batches) */
  int march_fuse;     /* 1: run pairs of smoothing sweeps in one launch, on
                         marching and square-tile levels alike (two sweeps
                         per read of x and b; the result is unchanged);
                         0: one sweep per launch */
  int autotune;       /* 1: at the first solve, time the square tiles and
                         row-marching tiles of 32..256 rows on every level
                         for both smoothers over the batch, and use the
                         fastest (overrides march_min / march_rows; the
                         arithmetic, hence the result, is the same) */
  int graphs;         /* 1: capture each block step's preconditioner (the
                         gather of the defects, prec_steps edge V-cycles and
                         the lifting with its node V-cycles) into a HIP graph
                         and replay it; it is captured again only when the
                         number of preconditioned vectors changes.  Same
                         kernels and arguments, hence the same result.
                         Ignored with phase_kernels or smoother_steps 0,
                         and by MX_METHOD_LOBPCG. */
  int method;         /* MX_METHOD_PINVIT (0): block PINVIT, Rayleigh-Ritz
                         over the preconditioned block (exp_bloch2d.c:
                         956-1053); MX_METHOD_LOBPCG: Rayleigh-Ritz over the
                         block, its preconditioned update and the previous
                         step's direction, 3 block columns (the reference's
                         LOBPCG program, test_lobpcg2d.c:247-312; block <= 16,
                         lock = 1: a converged pair's update direction is
                         rounding noise, which the three-block basis would
                         mix back into it).  Same preconditioner, same
                         convergence test. */
} mx_config;

typedef struct mx_solver mx_solver;

/* eta: (n0 << L)^2 values of 1/eps per cell, row-major (grid2d.h:33). */
int mx_solver_create(const mx_config *cfg, const double *eta, mx_solver **out);
void mx_solver_destroy(mx_solver *s);

typedef struct {
  int nprob;          /* problems in this batch, <= max_problems */
  const double *kx;   /* nprob Bloch parameters: xf = e^{i kx}, yf = e^{i ky} */
  const double *ky;
  int neig;           /* eigenpairs that must converge per problem (<= block) */
  double tol;         /* relative residual |K e - l M e| / (|l| |M e|) */
  int maxit;          /* block steps per problem at most */
  unsigned long long seed; /* start block when E0 is null: complex Gaussian */
  const double *E0;   /* optional host start block: nprob*block vectors */
} mx_batch;

typedef struct {
  double *lambda;     /* nprob*block Ritz values of K */
  double *relres;     /* nprob*block relative residuals */
  double *absres;     /* nprob*block |(K + (1 - l') M) e|_2, l' = l + 1 (the
                         reference's defect norm, exp_bloch2d.c:1013-1018) */
  int *iters;         /* nprob block steps to convergence (or maxit) */
  int *converged;     /* nprob flags */
} mx_result;

/* Batched block PINVIT + Rayleigh-Ritz: the exp_bloch2d k-loop body for
   nprob independent Bloch parameters at once (bloch_batch_solve). */
int mx_solve(mx_solver *s, const mx_batch *b, mx_result *r);

/* Copy problem p's current Ritz vectors (block vectors) to host memory. */
int mx_get_vectors(mx_solver *s, int p, double *E);

/* Adjoint: grad[c] = sum_{i in [first, first+count)} d lambda_i / d eta_c for
   problem p of the last solve, i.e. |curl e_i|_c^2 / (hx hy) / (e_i^H M e_i)
   summed over a (degenerate) cluster.  grad: (n0 << L)^2 doubles. */
int mx_eta_gradient(mx_solver *s, int p, int first, int count, double *grad);

/* Rayleigh-Ritz extrapolation of start blocks along a Bloch path
   (rayleigh_extrapolation, exp_blochsurf.c:332-406; the first- and
   second-order cases of exp_bloch2d.c:825-894).  For problem p of b (only
   b->nprob, b->kx and b->ky are read), H holds sources * ieigs host vectors
   back to back (complex FP64 in mx_get_vectors' layout; the reference's
   sources are the current, old and very old eigenvector blocks).  On the
   device they are M-orthonormalised by Householder reflections at k_p, the
   Ritz matrix of K + M is formed over them and solved, and the lowest block
   Ritz vectors go to E0 (host, nprob * block vectors: mx_batch.E0 of the next
   mx_solve); lambda (optional, nprob * block) receives their Ritz values of
   K.  block <= sources * ieigs <= 48.  Errors: MX_EINVAL, MX_ENOMEM,
   MX_EDEVICE, MX_ENOTFINITE (a non-finite Ritz matrix entry). */
int mx_extrapolate(mx_solver *s, const mx_batch *b, int sources, int ieigs, const double *H, double *E0,
                   double *lambda);

typedef struct {
  long long block_steps;   /* PINVIT block steps, summed over problems */
  long long edge_vcycles;  /* V-cycles (each over all active vectors) */
  long long node_vcycles;
  long long launches;      /* kernel launches issued */
  long long active_vectors; /* preconditioned vectors, summed over steps */
  double solve_seconds;    /* wall time of the last mx_solve */
  long long device_bytes;  /* device memory held by the solver */
  long long graph_replays;  /* preconditioner graph launches */
  long long graph_captures; /* captures (new or updated graphs) */
} mx_stats;

int mx_get_stats(const mx_solver *s, mx_stats *st);

/* Smoother tiling per level 0..levels: 0 = square tiles, r > 0 = row-marching
   with r rows per workgroup (the autotuned choice when autotune is set and a
   solve has run, else the march_min / march_rows rule for a launch over all
   max_problems * block vectors). */
int mx_get_tuning(const mx_solver *s, int *edge_rows, int *node_rows);

/* Time `reps` launches of one kernel of the hot path on the library stream
   with HIP events, on the data of the last solve.  kind: 0 = one AFW colour
   phase on the finest level over the whole batch, 1 = one edge V(nu,nu)
   cycle, 2 = one node V(nu,nu) cycle, 3 = one node GS colour phase on the
   finest level, 4 = one tile-fused AFW sweep on the finest level, 5 = one
   tile-fused nodal sweep on the finest level, 6 = one row-marching AFW
   sweep, 7 = one row-marching nodal sweep, 8 = a plain device copy of the
   finest-level multigrid vectors (read + write, the reference point for
   counter-based traffic), 9 = the Gram products of every problem's block
   as the solve forms them (FP64; LOBPCG: the 3 x block vectors [E W P]),
   10 / 11 = kinds 6 / 7 with two sweeps in one pass (bytes counted
   per sweep, as for one sweep each), 12 / 13 / 14 = the two-sweep row-marching
   AFW passes the solve runs on the finest level (V(2,2) with march_fuse):
   12 = zero start, two forward sweeps, residual and restriction (the first
   V-cycle's way down), 13 = the same without the zero start (later cycles),
   14 = prolongation of the coarse correction and two backward sweeps (every
   cycle's way up), with the solve's tiling for the whole batch, 15 / 16 / 17 =
   the same three passes of the nodal smoother (the lifting's node V-cycles);
   bytes count each fused part as its own stream.  avg_ms: mean duration of one launch
   (cycle); bytes: algorithmic HBM bytes of one launch (cycle). */
int mx_time_kernel(mx_solver *s, int kind, int reps, double *avg_ms, double *bytes);

/* Alternative names of the same entry points. */
#define bloch_batch_solve mx_solve
#define adjoint_eta_gradient mx_eta_gradient

#ifdef __cplusplus
}
#endif

#endif
