/* glr.h — GLR build and the downstream operations, MI355X engine.
 *
 * Drop-in for the reference's include/lgpsf_hessian/glr.h:50-169 (same
 * names, signatures, enum values, struct field order and defaults):
 *   F = M^{1/2} Z^{-1} B Z^{-1} M^{1/2} ~= U diag(lam) U^T   (build)
 *   f(F + cI) = U [f(lam + c) - f(c)] U^T + f(c) I           (operations)
 * with H(c) = B + c R.  Communicator size 1; B and Z are AIJ (or DENSE)
 * Mats of the PETSc subset; U, lam, B, Z and the AMG hierarchy stay
 * resident in HBM.  The stage-2.5 deflation correction against the true
 * Hessian (reference glr.h:170-270) is declared below with the same
 * structs and defaults. */
#ifndef LGPSF_HESSIAN_GLR_H
#define LGPSF_HESSIAN_GLR_H

#include <petscmat.h>
#include <petscvec.h>
#include "lgpsf_hessian/prior.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lgh_glr lgh_glr_t;

typedef enum
{
  LGH_GLR_DEFAULT    = 0,   /* resolves to REPLICATED                     */
  LGH_GLR_REPLICATED = 1,   /* explicit U = Q C (n x kept) in HBM          */
  LGH_GLR_SCALAPACK  = 2    /* the reference's at-scale backend at
                               communicator size 1 (a 1 x 1 process grid):
                               implicit U = Q V (Q n x ell, V ell x kept;
                               U never formed), BCGS2 orthonormalization,
                               check = eigenvalue cross-check of two
                               eigensolvers instead of the residual sweep
                               (glr_scalapack_impl.h:505-688) */
}
lgh_glr_backend_t;

typedef enum
{
  LGH_TREAT_FLIP = 0,
  LGH_TREAT_RELU = 1,
  LGH_TREAT_NONE = 2
}
lgh_treat_t;

typedef struct lgh_glr_opts
{
  int               ell;
  int               q_power;
  double            trunc_abs;
  double            trunc_rel;
  lgh_treat_t       treat;
  unsigned long     seed;
  lgh_glr_backend_t backend;
  int               nb;
  int               panel;
  int               grid2d;
  int               check;
}
lgh_glr_opts_t;

lgh_glr_opts_t lgh_glr_opts_default (void);

typedef struct lgh_glr_report
{
  int    kept;
  double lam_abs_max;
  double lam_raw_min, lam_raw_max;
  double next_abs;
  int    n_negative_raw;
  double resid_max;
  double t_operator, t_dense;
}
lgh_glr_report_t;

int lgh_glr_compute (Mat B, lgh_prior_t *prior, const lgh_glr_opts_t *opts,
                     lgh_glr_t **glr, lgh_glr_report_t *report);
/* incremental: hashed columns ell..ell+k_new-1, power passes on the new
 * block, BCGS2 against the kept Q, bordered projection, re-eigensolve */
int lgh_glr_extend (lgh_glr_t *glr, int k_new, lgh_glr_report_t *report);
void lgh_glr_destroy (lgh_glr_t *glr);

int lgh_glr_solve (lgh_glr_t *glr, double c, Vec rhs, Vec x);
int lgh_glr_apply (lgh_glr_t *glr, double c, Vec v, Vec y);
int lgh_glr_sample (lgh_glr_t *glr, double c, Vec xi, Vec x);

typedef enum
{
  LGH_FACTOR_G     = 0,
  LGH_FACTOR_GT    = 1,
  LGH_FACTOR_GINV  = 2,
  LGH_FACTOR_GINVT = 3
}
lgh_factor_t;

int lgh_glr_factor (lgh_glr_t *glr, double c, lgh_factor_t which,
                    Vec in, Vec out);
double lgh_glr_logdet (lgh_glr_t *glr, double c);
int lgh_glr_pow (lgh_glr_t *glr, double c, double p, Vec in, Vec out);
int lgh_glr_filter (lgh_glr_t *glr, const double *f_lam, double f_c,
                    Vec in, Vec out);
int lgh_glr_eigs (const lgh_glr_t *glr, int *nkept, const double **lam);

/* ---- stage 2.5: deflation correction against the true Hessian --------
 * Replaces glr.h:170-270 of the reference.  The whitened error operator
 *     E_w = S^- (F_true + c0 I) S^- - I,   S^- = (U lam U^T + c0 I)^{-1/2},
 *     F_true = M^{1/2} Z^{-1} H_d Z^{-T} M^{1/2},
 * is probed, its dominant eigenpairs (V, d) priced exactly with true-
 * Hessian applies, and U lam U^T + (S^+ V) d (S^+ V)^T folded into a new
 * (U', lam') kept SIGNED.  Afterwards every operation serves the
 * corrected operator at every shift c > floor = max(0, -min lam'),
 * lgh_glr_apply switches to the corrected low-rank action, and extend and
 * a second correction are refused. */
typedef struct lgh_glr_correct_opts
{
  double        c0;         /* reference shift of the error metric (1.0)  */
  int           rank;       /* max correction rank; -1 = all              */
  int           applies;    /* true-Hessian apply budget (30)             */
  int           n_qc;       /* probes path: trailing pairs held out (0)   */
  int           q_power;    /* fresh path: extra power passes (0)         */
  double        clamp_eps;  /* d clamped to >= -1 + clamp_eps (0.05)      */
  double        d_keep_min; /* drop priced modes with |d| below this (0)  */
  double        drop_tol;   /* relative basis drop tolerance (1e-12)      */
  unsigned long seed;       /* fresh-path hashed sketch seed (424242)     */
  int           verbose;
}
lgh_glr_correct_opts_t;

lgh_glr_correct_opts_t lgh_glr_correct_opts_default (void);

typedef struct lgh_glr_correct_report
{
  int    kept;              /* correction modes grafted                   */
  int    basis;             /* error-basis dimension available            */
  int    clamped;           /* Rayleigh values clamped at -1 + clamp_eps  */
  int    hd_applies;        /* true-Hessian applies spent                 */
  double d_min, d_max;      /* correction eigenvalues, post-clamp         */
  double floor;             /* shift validity: operations need c > floor  */
  double qc_before, qc_after;   /* whitened energy ratio on held-out
                                   probes; -1 when unavailable            */
  double t_operator, t_dense;
}
lgh_glr_correct_report_t;

/* y = H_d x on host arrays of length n (communicator size 1) */
#ifndef LGH_HESSIAN_FN_DEFINED
#define LGH_HESSIAN_FN_DEFINED
typedef void (*lgh_hessian_fn) (const double *in_local, double *out_local,
                                void *ctx);
#endif

/* value pass: the error basis from k probe pairs (V, HV = H_d V, n x k
 * column-major host arrays), the apply budget spent on exact Rayleigh
 * values (replaces glr.h:248-251) */
int lgh_glr_correct_probes (lgh_glr_t *glr, int k, const double *V,
                            const double *HV, lgh_hessian_fn hessian_apply,
                            void *ctx, const lgh_glr_correct_opts_t *opts,
                            lgh_glr_correct_report_t *report);

/* fresh randomized correction from a hashed Gaussian sketch of E_w
 * (replaces glr.h:256-259) */
int lgh_glr_correct (lgh_glr_t *glr, lgh_hessian_fn hessian_apply, void *ctx,
                     const lgh_glr_correct_opts_t *opts,
                     lgh_glr_correct_report_t *report);

#ifdef __cplusplus
}
#endif
#endif
