/* prior.h — the prior R = Z M^{-1} Z (M lumped/diagonal, Z symmetric), as
 * the GLR engine consumes it.  Same types, functions, struct layouts and
 * defaults as the reference's include/lgpsf_hessian/prior.h:38-110; this
 * build's Z-solve machinery runs on the GPU (see zsolve tiers below). */
#ifndef LGPSF_HESSIAN_PRIOR_H
#define LGPSF_HESSIAN_PRIOR_H

#include <petscmat.h>
#include <petscvec.h>
#include <petscksp.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lgh_prior lgh_prior_t;

typedef void (*lgh_vec_fn) (Vec x, Vec y, void *ctx);
typedef void (*lgh_mat_fn) (Mat X, Mat Y, void *ctx);

/* replaces prior.h:47-57 */
typedef struct lgh_prior_callbacks
{
  lgh_vec_fn applyZ;
  lgh_vec_fn applyZt;          /* NULL: Z symmetric                        */
  lgh_vec_fn solveZ;
  lgh_vec_fn solveZt;          /* NULL: Z symmetric                        */
  lgh_mat_fn solveZ_blocked;   /* optional (host MATDENSE blocks)          */
  lgh_mat_fn solveZt_blocked;
  void      *ctx;
}
lgh_prior_callbacks_t;

/* replaces prior.h:61-63 */
int lgh_prior_create_callbacks (Vec mass_lumps,
                                const lgh_prior_callbacks_t *callbacks,
                                lgh_prior_t **prior);

/* replaces prior.h:76-91.  blocked_mode:
 *   0 — per-column solves (block PCG, V-cycle preconditioner, ksp_rtol);
 *   2, 3 — blocked Chebyshev around the AMG V-cycle (host-built hierarchy,
 *          applied on the GPU; 2 and 3 are the same tier in this engine);
 *   4 — blocked Chebyshev with the Jacobi preconditioner, one fused kernel
 *       per iteration (extension; no reference counterpart).
 * Vector solves (solve / sample / factor) always use the PCG path.       */
typedef struct lgh_prior_mat_opts
{
  double ksp_rtol;
  int    blocked_mode;
  int    tile;
  double cheb_rtol;
  int    smax_krylov;
  double smoother_top;
  int    nu_force;
  int    verbose;
}
lgh_prior_mat_opts_t;

lgh_prior_mat_opts_t lgh_prior_mat_opts_default (void);

/* replaces prior.h:94-96 */
int lgh_prior_create_mat (Mat Z, Vec mass_lumps,
                          const lgh_prior_mat_opts_t *opts,
                          lgh_prior_t **prior);

/* replaces prior.h:103-105.  The PETSc subset has no KSP: returns
 * PETSC_ERR_SUP. */
int lgh_prior_create_ksp (KSP ksp, Vec mass_lumps,
                          const lgh_prior_mat_opts_t *opts,
                          lgh_prior_t **prior);

void lgh_prior_destroy (lgh_prior_t *prior);

/* y = Z M^{-1} Z^T x  (replaces prior.h:110) */
int lgh_prior_apply (lgh_prior_t *prior, Vec x, Vec y);

#ifdef __cplusplus
}
#endif
#endif
