/* lgh_ext.h — extensions of the MI355X engine beyond the reference API
 * (no counterpart in the reference): precision policy of the build,
 * batched multi-shift downstream operations, and read-only introspection
 * used by the parity tests and the benchmark. */
#ifndef LGPSF_HESSIAN_LGH_EXT_H
#define LGPSF_HESSIAN_LGH_EXT_H

#include <stdint.h>

#include "lgpsf_hessian/glr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LGH_PREC_F64  0
#define LGH_PREC_F32  1
#define LGH_PREC_BF16 2   /* storage only (fp32 arithmetic); V-cycle only */

/* Precision policy of the build.  All LGH_PREC_F64 reproduces the
 * reference arithmetic ("plain").  The mixed policy (lgh_glr_xopts_mixed)
 * runs every sweep of the build with FP64 vectors and B inside an FP64
 * Chebyshev iteration whose AMG V-cycle stores its vectors in BF16 (FP32
 * matrices and arithmetic; the Chebyshev step rounds each new direction
 * once and updates X and R with that same value, so the Z solves stay
 * FP64-accurate), keeps CholQR2, T = Q^T F Q, the eigensolve and U in
 * FP64, runs the residual-gate sweep with FP32 vectors around a BF16
 * V-cycle (a diagnostic at FP32 rounding either way), and takes downstream
 * calls (one column or many) through the same blocked tier.
 * prec_sketch = F32 runs the sketch sweep with FP32 vectors, B and Z
 * instead (its V-cycle in prec_sketch_precond).  A V-cycle precision
 * above its sweep's vector precision means "the vectors' own". */
typedef struct lgh_glr_xopts
{
  int prec_sketch;          /* sketch sweep Y = F Omega: F64 or F32        */
  int prec_precond;         /* V-cycle inside FP64-vector sweeps:
                               F64, F32 or BF16                            */
  int prec_check;           /* residual-gate sweep (opts.check): F64, F32  */
  int prec_sketch_precond;  /* V-cycle inside FP32-vector sweeps (sketch,
                               check): F32 (or F64) or BF16                */
  int solve_blocked_min;    /* downstream solve / sample / factor calls on
                               at least this many columns use the blocked
                               tier (FP64 Chebyshev around the prec_precond
                               V-cycle, 64-column panels) for their Z
                               solves instead of per-column PCG to
                               ksp_rtol; 0 = never (the reference path)    */
  int sweep_panels;         /* 64-column panels per F sweep chunk (the
                               Z-solve work of a sweep grows with the panels
                               solved together); 0 = all panels at once
                               unless that would not fit in device memory
                               beside the sketch blocks.  Results do not
                               depend on it (panels are independent)       */
}
lgh_glr_xopts_t;

lgh_glr_xopts_t lgh_glr_xopts_default (void);   /* all LGH_PREC_F64 */
lgh_glr_xopts_t lgh_glr_xopts_mixed (void);

int lgh_glr_compute_x (Mat B, lgh_prior_t *prior, const lgh_glr_opts_t *opts,
                       const lgh_glr_xopts_t *xopts, lgh_glr_t **glr,
                       lgh_glr_report_t *report);

/* Batched operations on m host columns (column-major, leading dimension
 * ld), column j at shift c[j]:  x_j = H(c_j)^{-1} rhs_j  and
 * x_j = G(c_j)^{-T} xi_j.  Same arithmetic per column as lgh_glr_solve /
 * lgh_glr_sample. */
int lgh_glr_solve_block (lgh_glr_t *glr, int m, const double *c,
                         const double *rhs, int ldr, double *x, int ldx);
int lgh_glr_sample_block (lgh_glr_t *glr, int m, const double *c,
                          const double *xi, int ldr, double *x, int ldx);

/* The same two operations on blocks already in device memory, for callers
 * whose right-hand sides live in HBM (no host transfer): rhs / xi and x
 * are device pointers (column-major, leading dimensions >= n; host memory
 * is refused with PETSC_ERR_ARG_WRONG, a block that extends past its
 * device allocation with PETSC_ERR_ARG_SIZ), c a host array of m shifts.  The
 * engine reads the input on its own stream, so the work that wrote it must
 * have finished (synchronize the producing stream first); x is complete
 * when the call returns.  The same bits as lgh_glr_solve_block /
 * lgh_glr_sample_block on the same columns. */
int lgh_glr_solve_block_dev (lgh_glr_t *glr, int m, const double *c,
                             const double *rhs, int ldr, double *x, int ldx);
int lgh_glr_sample_block_dev (lgh_glr_t *glr, int m, const double *c,
                              const double *xi, int ldr, double *x, int ldx);

/* Batches of independent instances (no reference counterpart: the
 * reference builds one object per call).  Instance i runs the single
 * call with argument i (lgh_prior_create_mat, lgh_glr_compute_x): the
 * same kernels in the same order, on `workers` host threads (<= 0: 8),
 * each driving the GPU on its own stream, so the kernels of small
 * instances overlap.  Every kernel reduces in a fixed order and the
 * Cholesky and triangular inverse are the engine's own, so each instance's
 * result equals the single call's bit for bit (a block's partial last
 * panel is solved with the whole panels here, as narrow panels in the
 * single call: a column's bits do not depend on its panel's width).
 * Instances that share an object (Z, mass, B or prior) run one
 * after another on one thread.  opts: one per instance or NULL (defaults);
 * xopts: shared or NULL.  On failure every output is destroyed and set to
 * NULL, and the error names the lowest failing instance. */
int lgh_prior_create_mat_batch (int count, Mat *Z, Vec *mass,
                                const lgh_prior_mat_opts_t *opts, int workers,
                                lgh_prior_t **priors);
int lgh_glr_compute_batch (int count, Mat *B, lgh_prior_t **priors,
                           const lgh_glr_opts_t *opts,
                           const lgh_glr_xopts_t *xopts, int workers,
                           lgh_glr_t **glrs, lgh_glr_report_t *reports);

typedef struct lgh_prior_info
{
  int    is_mat;        /* 1: assembled Z, 0: callbacks                    */
  int    blocked_mode;
  int    nit;           /* blocked Chebyshev iterations                    */
  double lmin, lmax;    /* margined preconditioned interval                */
  int    levels;        /* AMG levels (0 when none)                        */
  int    n_coarse;      /* coarsest AMG size                               */
  int    bound_its;     /* iterations of the interval estimate             */
  int    cycle;         /* AMG cycle index: 1 V-cycle, 2 W-cycle           */
}
lgh_prior_info_t;

int lgh_prior_get_info (const lgh_prior_t *prior, lgh_prior_info_t *info);

/* Z-solves through the prior's tiers (host column-major blocks):
 * which = 0 blocked tier (what the build uses), 1 vector (PCG) path;
 * prec = 0 FP64, 1 FP32, 2 FP64 vectors around the FP32 V-cycle,
 * 3 FP64 vectors around the BF16 V-cycle, 4 FP32 around the BF16 V-cycle. */
int lgh_prior_solve_host (lgh_prior_t *prior, int which, int prec, int m,
                          const double *X, int ldx, double *Y, int ldy);

/* Request loop over one resident GLR: callers on any thread submit solve /
 * sample / logdet requests (shift c, m right-hand sides, column-major host
 * blocks that stay valid until the wait returns); batches of the oldest
 * request's operation are formed from everything queued (each column with
 * its own shift) into pinned staging blocks (two per device thread), so
 * requests that arrive after a batch is formed join the next one.  One
 * thread copies a batch's columns in, device_streams threads run batches
 * on the device at once (each on its own streams), one copies the results
 * out: the copies of one batch overlap the device work of the others.  The
 * batch width is capped by a device workspace
 * budget; submit blocks while more than queue_cols columns are queued.
 * A request's result does not depend on what it was batched with.  While a
 * server exists, the GLR must not be used or destroyed by other calls.
 * (No reference counterpart: the reference's operations are synchronous,
 * one Vec at a time.) */
typedef struct lgh_serve lgh_serve_t;

#define LGH_SERVE_SOLVE  0   /* x = H(c)^{-1} rhs          (lgh_glr_solve)  */
#define LGH_SERVE_SAMPLE 1   /* x = G(c)^{-T} xi           (lgh_glr_sample) */
#define LGH_SERVE_LOGDET 2   /* logdet(I + F / c)          (lgh_glr_logdet) */

typedef struct lgh_serve_opts
{
  int    max_batch_cols;    /* columns per device batch (multiple of 64)    */
  double device_budget_gb;  /* workspace budget of one batch; caps the width */
  int    queue_cols;        /* queued columns before submit blocks          */
  int    device_streams;    /* batches on the device at once, each on its own
                               device thread, streams and Z-solve buffers
                               (1 .. 8, default 3; the budget is per batch) */
}
lgh_serve_opts_t;

typedef struct lgh_serve_stats
{
  int64_t batches, columns, requests;
  int     max_batch_cols;   /* widest batch run so far                      */
  int     batch_cols;       /* the admitted batch width                      */
  double  busy_s;           /* device thread's time in batches              */
  double  stage_s;          /* host copies between the callers' columns and
                               the pinned staging blocks (staging threads,
                               overlapping the device work)                 */
  int     max_running;      /* most batches on the device at once so far    */
}
lgh_serve_stats_t;

lgh_serve_opts_t lgh_serve_opts_default (void);
int  lgh_serve_create (lgh_glr_t *glr, const lgh_serve_opts_t *opts,
                       lgh_serve_t **server);
/* m, in, ldi, out, ldo are ignored for LGH_SERVE_LOGDET */
int  lgh_serve_submit (lgh_serve_t *server, int op, double c, int m,
                       const double *in, int ldi, double *out, int ldo,
                       int64_t *ticket);
/* blocks until the request is done; logdet (LGH_SERVE_LOGDET) and the
 * submit-to-done latency in seconds are returned when non-NULL */
int  lgh_serve_wait (lgh_serve_t *server, int64_t ticket, double *logdet,
                     double *latency_s);
int  lgh_serve_get_stats (lgh_serve_t *server, lgh_serve_stats_t *stats);
void lgh_serve_destroy (lgh_serve_t *server);

/* B from the fit stage's row export (the CSR of lgh_fit_get_rows,
 * reference fit.h:160-165: row pointers, 64-bit global column ids, values;
 * communicator size 1) as an assembled AIJ Mat in one pass, replacing the
 * per-entry lgh_fit_get_mat assembly (reference glr.h:274-282).  Column
 * ids are range-checked, each row sorted by column with duplicates summed;
 * with sym_tol >= 0 the matrix must equal its transpose to
 * |b_ij - b_ji| <= sym_tol * max|b| (PETSC_ERR_ARG_WRONG otherwise). */
int lgh_mat_from_fit_rows (int n, const int *rowptr, const long *colgids,
                           const double *vals, double sym_tol, Mat *B);

/* copy of the kept U (n x kept, column-major) for tests */
int lgh_glr_get_U (const lgh_glr_t *glr, double *U, int ldu);

#ifdef __cplusplus
}
#endif
#endif
