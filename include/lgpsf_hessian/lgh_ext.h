/* lgh_ext.h — extensions of the MI355X engine beyond the reference API
 * (no counterpart in the reference): precision policy of the build,
 * batched multi-shift downstream operations, and read-only introspection
 * used by the parity tests and the benchmark. */
#ifndef LGPSF_HESSIAN_LGH_EXT_H
#define LGPSF_HESSIAN_LGH_EXT_H

#include "lgpsf_hessian/glr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LGH_PREC_F64  0
#define LGH_PREC_F32  1
#define LGH_PREC_BF16 2   /* storage only (fp32 arithmetic); V-cycle only */

/* Precision policy of the build.  All LGH_PREC_F64 reproduces the
 * reference arithmetic ("plain").  The mixed policy (lgh_glr_xopts_mixed)
 * runs the sketch sweep Y = F Omega with FP32 vectors, B, Z and V-cycle
 * (its rounding acts like a perturbation of the Gaussian sketch), runs
 * every later sweep with FP64 vectors inside an FP64 Chebyshev iteration
 * whose AMG V-cycle stores its vectors in BF16 (the Chebyshev step rounds
 * each new direction once and updates X and R with that same value, so the
 * Z solves stay FP64-accurate), keeps CholQR2, T = Q^T F Q, the
 * eigensolve and U in FP64, and runs the residual-gate sweep in FP32.
 * A V-cycle precision above its sweep's vector precision means "the
 * vectors' own". */
typedef struct lgh_glr_xopts
{
  int prec_sketch;          /* sketch sweep Y = F Omega: F64 or F32        */
  int prec_precond;         /* V-cycle inside FP64-vector sweeps:
                               F64, F32 or BF16                            */
  int prec_check;           /* residual-gate sweep (opts.check): F64, F32  */
  int prec_sketch_precond;  /* V-cycle inside FP32-vector sweeps (sketch,
                               check): F32 (or F64) or BF16                */
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

typedef struct lgh_prior_info
{
  int    is_mat;        /* 1: assembled Z, 0: callbacks                    */
  int    blocked_mode;
  int    nit;           /* blocked Chebyshev iterations                    */
  double lmin, lmax;    /* margined preconditioned interval                */
  int    levels;        /* AMG levels (0 when none)                        */
  int    n_coarse;      /* coarsest AMG size                               */
  int    bound_its;     /* iterations of the interval estimate             */
}
lgh_prior_info_t;

int lgh_prior_get_info (const lgh_prior_t *prior, lgh_prior_info_t *info);

/* Z-solves through the prior's tiers (host column-major blocks):
 * which = 0 blocked tier (what the build uses), 1 vector (PCG) path;
 * prec = 0 FP64, 1 FP32, 2 FP64 vectors around the FP32 V-cycle,
 * 3 FP64 vectors around the BF16 V-cycle, 4 FP32 around the BF16 V-cycle. */
int lgh_prior_solve_host (lgh_prior_t *prior, int which, int prec, int m,
                          const double *X, int ldx, double *Y, int ldy);

/* copy of the kept U (n x kept, column-major) for tests */
int lgh_glr_get_U (const lgh_glr_t *glr, double *U, int ldu);

#ifdef __cplusplus
}
#endif
#endif
