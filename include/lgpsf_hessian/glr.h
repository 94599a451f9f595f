/* glr.h — GLR build and the downstream operations, MI355X engine.
 *
 * Drop-in for the reference's include/lgpsf_hessian/glr.h:50-169 (same
 * names, signatures, enum values, struct field order and defaults):
 *   F = M^{1/2} Z^{-1} B Z^{-1} M^{1/2} ~= U diag(lam) U^T   (build)
 *   f(F + cI) = U [f(lam + c) - f(c)] U^T + f(c) I           (operations)
 * with H(c) = B + c R.  Communicator size 1; B and Z are AIJ (or DENSE)
 * Mats of the PETSc subset; U, lam, B, Z and the AMG hierarchy stay
 * resident in HBM.  The stage-2.5 correction (glr.h:200-270 of the
 * reference) is not part of this build. */
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
  LGH_GLR_DEFAULT    = 0,
  LGH_GLR_REPLICATED = 1,
  LGH_GLR_SCALAPACK  = 2    /* not in this build: PETSC_ERR_SUP */
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

#ifdef __cplusplus
}
#endif
#endif
