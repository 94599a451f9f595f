// engine.hpp — object layouts of lgh_prior / lgh_glr and the internal
// device-side operations the API functions compose.
#pragma once
#include <lgpsf_hessian/lgpsf_hessian.h>

#include "lgh_dev.hpp"
#include "zsolve.hpp"
#include "petsc_impl.hpp"

#include <mutex>
#include <vector>

struct lgh_prior {
  int refs = 1;
  Vec mass = nullptr, msqrt = nullptr, minvsqrt = nullptr, work = nullptr;
  double *d_msqrt = nullptr, *d_minvsqrt = nullptr, *d_minv = nullptr;
  lgh_prior_callbacks_t cb {};
  int is_mat = 0;
  // Mat path
  Mat Z = nullptr;
  lgh_prior_mat_opts_t o {};
  lgh::AmgHierarchy amg;
  int amg_ok = 0;
  lgh::DevMG *mg[3] = { nullptr, nullptr, nullptr };   // per lgh::Prec
  double *d_zdinv = nullptr;            // fp64 1/diag(Z) (Jacobi PCG)
  int nit = 0, bound_its = 0;
  double lmin = 0, lmax = 0;
  lgh::ChebWork cw[2];
  lgh::ChebMixedWork cwm;
  int n = 0;
  std::recursive_mutex lazy_mu;         // first-use device set-up
};

struct lgh_glr {
  Mat B = nullptr;
  lgh_prior_t *prior = nullptr;
  lgh_glr_opts_t opts {};
  lgh_glr_xopts_t x {};
  int n = 0;
  lgh::PMat U;                          // n x kept, fp64, W = 64
  std::vector<double> lam;
  int kept = 0;
  // implicit representation (backend LGH_GLR_SCALAPACK): U = Q V with V the
  // kept Ritz vectors (ell x kept, column-major) on the host and, zero-
  // padded to 64-multiples (ld = npanels(ell) * 64), on the device; U empty
  int implicit = 0;
  std::vector<double> V;
  double *dV = nullptr;
  // incremental-extension state: orthonormal Q (n x ell) and raw Q^T F Q
  lgh::PMat Q;
  std::vector<double> Araw;
  int ell = 0;
  int corrected = 0;
  double floor = 0.0;
};

namespace lgh {
// device-side prior operations on panel matrices (precision = X.prec)
int prior_setup_device (lgh_prior_t *p, int prec);
// trans: 1 selects the transposed callbacks (Z symmetric on the Mat path)
// pcp: V-cycle precision when lower than X's (F32 or BF16 around FP64 X:
// the mixed Chebyshev; BF16 around FP32 X); -1: X's precision.
// yscale (n values or null): Y is returned as diag(yscale) Z^{-1} X, the
// bits of k_diag_scale after the solve (the mixed Chebyshev writes it from
// its last step)
int prior_solve_block (lgh_prior_t *p, int trans, const PMat &X, PMat &Y,
                       int pcp = -1, const double *yscale = nullptr);
int prior_solve_vec (lgh_prior_t *p, int trans, const PMat &X, PMat &Y);
int prior_apply_Z (lgh_prior_t *p, int trans, const PMat &X, PMat &Y);
void prior_ref (lgh_prior_t *p);
// batched solve (op 0) or sample (op 3) of m host columns, column j at
// shift c[j]; tier: -1 by the GLR's policy, 0 per-column PCG, 1 blocked
int glr_run_block (lgh_glr_t *g, int op, int m, const double *c,
                   const double *in, int ldi, double *out, int ldo, int tier);
// Y = f_c X + U diag(f_lam - f_c) U^T X on m columns of W = 64 FP64 panels
// (the same filter for every column)
int glr_filter_panels (lgh_glr_t *g, const double *f_lam, double f_c,
                       const PMat &X, int m, PMat &Y);
// an implicit object's U = Q V formed in HBM (the object becomes explicit;
// no-op for an explicit one)
int glr_materialize_U (lgh_glr_t *g);
}
