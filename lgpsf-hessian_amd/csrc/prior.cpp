// prior.cpp — the prior object: lgh_prior_create_mat / _callbacks /
// _apply / _destroy (glr_common_impl.h:99-244, zsolve_impl.h:720-823) with
// the Z-solve tiers on the GPU and the AMG hierarchy built on the host.
#include "engine.hpp"

#include <mutex>

#include <cmath>
#include <cstring>

using namespace lgh;

extern "C" lgh_prior_mat_opts_t
lgh_prior_mat_opts_default (void)
{
  // defaults of glr_common_impl.h:99-112
  lgh_prior_mat_opts_t o;
  o.ksp_rtol = 1e-12;
  o.blocked_mode = 0;
  o.tile = 64;
  o.cheb_rtol = 1e-10;
  o.smax_krylov = 1;
  o.smoother_top = 1.1;
  o.nu_force = 0;
  o.verbose = 0;
  return o;
}

// glr_common_impl.h:114-132: msqrt = sqrt|m|, minvsqrt = 1/msqrt
static int
prior_init_common (Vec mass, lgh_prior_t **out)
{
  lgh_prior_t *p = new lgh_prior_t ();
  p->mass = mass;
  PetscObjectReference ((PetscObject) mass);
  VecDuplicate (mass, &p->msqrt);
  VecCopy (mass, p->msqrt);
  VecSqrtAbs (p->msqrt);
  VecDuplicate (p->msqrt, &p->minvsqrt);
  VecCopy (p->msqrt, p->minvsqrt);
  VecReciprocal (p->minvsqrt);
  VecDuplicate (mass, &p->work);
  p->n = mass->n;
  *out = p;
  return 0;
}

static int
upload_vecs (lgh_prior_t *p)
{
  if (p->d_msqrt) return 0;
  const size_t b = sizeof (double) * (size_t) p->n;
  p->d_msqrt = (double *) dmalloc (b);
  p->d_minvsqrt = (double *) dmalloc (b);
  p->d_minv = (double *) dmalloc (b);
  if (!p->d_msqrt || !p->d_minvsqrt || !p->d_minv) return 5;
  std::vector<double> minv ((size_t) p->n);
  for (int i = 0; i < p->n; i++) minv[i] = 1.0 / p->mass->a[i];
  LGH_TRY (h2d (p->d_msqrt, p->msqrt->a, b));
  LGH_TRY (h2d (p->d_minvsqrt, p->minvsqrt->a, b));
  LGH_TRY (h2d (p->d_minv, minv.data (), b));
  return 0;
}

extern "C" int
lgh_prior_create_callbacks (Vec mass, const lgh_prior_callbacks_t *cb,
                            lgh_prior_t **prior)
{
  PetscCheck (cb != NULL && cb->applyZ != NULL && cb->solveZ != NULL, 0,
              PETSC_ERR_ARG_NULL,
              "lgh_prior_create_callbacks: applyZ and solveZ are required");
  PetscCheck (mass != NULL, 0, PETSC_ERR_ARG_NULL,
              "lgh_prior_create_callbacks: mass is NULL");
  prior_init_common (mass, prior);
  (*prior)->cb = *cb;
  return 0;
}

extern "C" int
lgh_prior_create_ksp (KSP, Vec, const lgh_prior_mat_opts_t *, lgh_prior_t **)
{
  return lgh_petsc_error (PETSC_ERR_SUP,
                          "lgh_prior_create_ksp: the engine has no KSP; use "
                          "lgh_prior_create_mat or lgh_prior_create_callbacks");
}

// device mirror of the Mat-path machinery at one precision.  The device
// copies (vectors, Z, hierarchies) are made on first use; threads that
// solve with one prior at once (the serving loop's device threads, each
// with its own SolveWork) make them once, under the prior's lock.
int
lgh::prior_setup_device (lgh_prior_t *p, int prec)
{
  std::lock_guard<std::recursive_mutex> lk (p->lazy_mu);
  LGH_TRY (dev_init ());
  LGH_TRY (upload_vecs (p));
  if (!p->is_mat) return 0;
  DevCSR *dz = nullptr;
  LGH_TRY (mat_device_csr (p->Z, prec, 1, &dz));
  if (p->amg_ok && !p->mg[prec]) LGH_TRY (mg_upload (p->amg, prec, &p->mg[prec]));
  if (!p->d_zdinv) {
    std::vector<double> d ((size_t) p->n, 0.0);
    const int *rp = p->Z->rp.data (), *ci = p->Z->ci.data ();
    const double *v = p->Z->v.data ();
    for (int i = 0; i < p->n; i++)
      for (int k = rp[i]; k < rp[i + 1]; k++) if (ci[k] == i) d[i] += v[k];
    for (auto &x : d) x = 1.0 / x;
    p->d_zdinv = (double *) dmalloc (sizeof (double) * p->n);
    if (!p->d_zdinv) return 5;
    LGH_TRY (h2d (p->d_zdinv, d.data (), sizeof (double) * p->n));
  }
  return 0;
}

extern "C" int
lgh_prior_create_mat (Mat Z, Vec mass, const lgh_prior_mat_opts_t *opts,
                      lgh_prior_t **prior)
{
  lgh_prior_mat_opts_t o = opts ? *opts : lgh_prior_mat_opts_default ();
  PetscCheck (Z != NULL && mass != NULL, 0, PETSC_ERR_ARG_NULL,
              "lgh_prior_create_mat: Z and mass are required");
  PetscCheck (Z->type == LGH_MAT_AIJ && Z->assembled, 0,
              PETSC_ERR_ARG_WRONGSTATE,
              "lgh_prior_create_mat: Z must be an assembled AIJ Mat");
  PetscCheck (Z->m == Z->n && Z->m == mass->n, 0, PETSC_ERR_ARG_SIZ,
              "lgh_prior_create_mat: Z (%d x %d) and mass (%d) disagree",
              Z->m, Z->n, mass->n);
  PetscCheck (o.blocked_mode == 0 || o.blocked_mode == 1
              || o.blocked_mode == 2 || o.blocked_mode == 3
              || o.blocked_mode == 4, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_prior_create_mat: blocked_mode %d unknown", o.blocked_mode);
  lgh_prior_t *p = nullptr;
  prior_init_common (mass, &p);
  p->is_mat = 1;
  p->Z = Z;
  PetscObjectReference ((PetscObject) Z);
  p->o = o;
  // the AMG hierarchy (the reference's PCGAMG, zsolve_impl.h:808-818)
  {
    HostCSR A;
    A.n = Z->m; A.m = Z->n; A.rp = Z->rp; A.ci = Z->ci; A.v = Z->v;
    AmgOpts ao;
    ao.smax_krylov = o.smax_krylov;
    ao.smoother_top = o.smoother_top;
    ao.nu_force = o.nu_force;
    p->amg_ok = (amg_setup (A, ao, p->amg) == 0);
  }
  int rc = prior_setup_device (p, F64);
  if (rc) {
    lgh_prior_destroy (p);
    return lgh_petsc_error (PETSC_ERR_LIB, "lgh_prior_create_mat: %s",
                            lgh::last_error ());
  }
  // interval estimate of the blocked tier (zsolve_impl.h:737-772): CG on
  // rhs = randn_at(0xB0BD5, i, 0), rtol 1e-12, maxit 200, preconditioned by
  // the V-cycle (modes 2/3) or Jacobi (mode 4); margins 0.9 / 1.1
  if (o.blocked_mode >= 2) {
    if (o.blocked_mode <= 3 && !p->amg_ok) {
      lgh_prior_destroy (p);
      return lgh_petsc_error (PETSC_ERR_LIB,
                              "lgh_prior_create_mat: AMG setup failed "
                              "(coarsening stalled above 2000 unknowns)");
    }
    // Lanczos extremes of the preconditioned Z from one PCG run
    auto estimate = [&] (double *emin, double *emax, int *its) -> int {
      PMat b, x;
      PcgInfo info;
      int rc2 = pmat_alloc (b, p->n, 1, 1, F64);
      if (!rc2) rc2 = pmat_alloc (x, p->n, 1, 1, F64);
      if (!rc2) rc2 = k_randn (b, 0xB0BD5ULL, 0, 0);
      DevCSR *dz = nullptr;
      if (!rc2) rc2 = mat_device_csr (Z, F64, 1, &dz);
      if (!rc2)
        rc2 = pcg_solve (*dz, o.blocked_mode <= 3 ? p->mg[F64] : nullptr,
                         p->d_zdinv, b, x, 1e-12, 200, &info);
      pmat_free (b);
      pmat_free (x);
      if (rc2) return rc2;
      if (info.its == 0) { set_error ("no PCG iteration"); return 9; }
      const int it = info.its;
      std::vector<double> d (it), e (it);
      for (int k = 0; k < it; k++) {
        d[k] = 1.0 / info.alpha[k] + (k > 0 ? info.beta[k - 1] / info.alpha[k - 1] : 0.0);
        e[k] = std::sqrt (info.beta[k]) / info.alpha[k];
      }
      tridiag_extremes (it, d.data (), e.data (), emin, emax);
      *its = it;
      return 0;
    };
    // zsolve_impl.h:93-102
    auto cheb_count = [&] (double emin, double emax) {
      const double kap = (1.1 * emax) / (0.9 * emin);
      const double rho = (std::sqrt (kap) - 1.) / (std::sqrt (kap) + 1.);
      const int nit = (int) std::ceil (std::log (2. / o.cheb_rtol) / std::log (1. / rho));
      return nit < 1 ? 1 : nit;
    };
    double emin = 0, emax = 0;
    int it = 0;
    rc = estimate (&emin, &emax, &it);
    // cycle index: the W-cycle when its smaller Chebyshev count pays for
    // its extra coarse-level work by at least 10 % (bytes per column per
    // Chebyshev iteration: the fused outer step, ~40 B a row + Z's CSR,
    // plus one cycle); deterministic, from the hierarchy's sizes only
    if (!rc && o.blocked_mode <= 3 && p->amg.lev.size () >= 3) {
      double emin2 = 0, emax2 = 0;
      int it2 = 0;
      p->mg[F64]->gamma = 2;
      rc = estimate (&emin2, &emax2, &it2);
      p->mg[F64]->gamma = 1;
      if (!rc) {
        const double step = 40.0 * p->n + 12.0 * (double) Z->rp[Z->m] / 64.0;
        const double c1 = cheb_count (emin, emax) * (step + amg_cycle_cost (p->amg, 1));
        const double c2 = cheb_count (emin2, emax2) * (step + amg_cycle_cost (p->amg, 2));
        if (o.verbose)
          PetscPrintf (0, "lgh_prior: V-cycle nit %d cost %.3e, W-cycle nit %d "
                       "cost %.3e\n", cheb_count (emin, emax), c1,
                       cheb_count (emin2, emax2), c2);
        if (c2 < 0.9 * c1) {
          p->amg.gamma = 2;
          for (auto *m : p->mg) if (m) m->gamma = 2;
          emin = emin2; emax = emax2; it = it2;
        }
      }
    }
    if (rc) {
      lgh_prior_destroy (p);
      return lgh_petsc_error (PETSC_ERR_LIB, "lgh_prior_create_mat: interval "
                              "estimate failed: %s", lgh::last_error ());
    }
    p->lmin = 0.9 * emin;
    p->lmax = 1.1 * emax;
    p->nit = cheb_count (emin, emax);
    p->bound_its = it;
    if (o.verbose)
      PetscPrintf (0, "lgh_prior: mode %d bounds [%.3e, %.3e] (%d its) -> "
                   "chebyshev nit %d\n", o.blocked_mode, emin, emax, it, p->nit);
  }
  *prior = p;
  return 0;
}

void
lgh::prior_ref (lgh_prior_t *p)
{
  p->refs++;
}

extern "C" void
lgh_prior_destroy (lgh_prior_t *p)
{
  if (p == NULL || --p->refs > 0) return;
  for (auto *m : p->mg) delete m;
  dfree (p->d_msqrt); dfree (p->d_minvsqrt); dfree (p->d_minv);
  dfree (p->d_zdinv);
  VecDestroy (&p->msqrt);
  VecDestroy (&p->minvsqrt);
  VecDestroy (&p->work);
  VecDestroy (&p->mass);
  MatDestroy (&p->Z);
  delete p;
}

// ---- device-side operations --------------------------------------------------

// host staging for callback priors: column loop over the user's Vec
// callbacks (or its MATDENSE blocked callback), as glr_common_impl.h:206-234
static int
cb_block (lgh_prior_t *p, int kind, const PMat &X, PMat &Y)
{
  // kind: 0 solveZ, 1 solveZt, 2 applyZ, 3 applyZt
  const int n = X.nrows, nc = X.ncols;
  std::vector<double> hx ((size_t) n * nc), hy ((size_t) n * nc);
  PMat tx, ty;
  LGH_TRY (pmat_alloc (tx, n, nc, 1, F64));
  LGH_TRY (pmat_alloc (ty, n, nc, 1, F64));
  int rc = k_copy_cols (X, 0, tx, 0, nc);
  if (!rc) rc = d2h (hx.data (), tx.p, sizeof (double) * n * nc);
  lgh_mat_fn blocked = nullptr;
  if (kind == 0) blocked = p->cb.solveZ_blocked;
  else if (kind == 1)
    blocked = p->cb.solveZt_blocked ? p->cb.solveZt_blocked
            : (p->cb.solveZt ? nullptr : p->cb.solveZ_blocked);
  if (!rc && blocked) {
    Mat MX, MY;
    MatCreateSeqDense (0, n, nc, hx.data (), &MX);
    MatCreateSeqDense (0, n, nc, hy.data (), &MY);
    blocked (MX, MY, p->cb.ctx);
    MatDestroy (&MX);
    MatDestroy (&MY);
  }
  else if (!rc) {
    for (int j = 0; j < nc; j++) {
      Vec vx, vy;
      VecCreateSeqWithArray (0, 1, n, hx.data () + (size_t) j * n, &vx);
      VecCreateSeqWithArray (0, 1, n, hy.data () + (size_t) j * n, &vy);
      lgh_vec_fn f;
      switch (kind) {
      case 0: f = p->cb.solveZ; break;
      case 1: f = p->cb.solveZt ? p->cb.solveZt : p->cb.solveZ; break;
      case 2: f = p->cb.applyZ; break;
      default: f = p->cb.applyZt ? p->cb.applyZt : p->cb.applyZ; break;
      }
      f (vx, vy, p->cb.ctx);
      VecDestroy (&vx);
      VecDestroy (&vy);
    }
  }
  if (!rc) rc = h2d (ty.p, hy.data (), sizeof (double) * n * nc);
  if (!rc) rc = k_copy_cols (ty, 0, Y, 0, nc);
  if (!rc) rc = dev_sync ();
  pmat_free (tx);
  pmat_free (ty);
  return rc;
}

int
lgh::prior_solve_block (lgh_prior_t *p, int trans, const PMat &X, PMat &Y,
                        int pcp, const double *yscale)
{
  // the other tiers scale after the solve
  auto scaled = [&] (int rc) -> int {
    if (rc || !yscale) return rc;
    return k_diag_scale (Y, yscale, Y);
  };
  if (!p->is_mat) return scaled (cb_block (p, trans ? 1 : 0, X, Y));
  const int mode = p->o.blocked_mode;
  if (mode < 2) return scaled (prior_solve_vec (p, trans, X, Y));
  // a V-cycle in a lower precision than the vectors (W = 64): FP32 or BF16
  // around FP64 vectors, BF16 around FP32 vectors
  const bool lower = (X.prec == F64 && (pcp == F32 || pcp == BF16))
                     || (X.prec == F32 && pcp == BF16);
  const bool mixed = mode <= 3 && p->amg_ok && lower && X.W == 64;
  DevCSR *dz = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk (p->lazy_mu);
    LGH_TRY (prior_setup_device (p, X.prec));
    LGH_TRY (mat_device_csr (p->Z, X.prec, 1, &dz));
    if (mixed && !p->mg[pcp]) LGH_TRY (mg_upload (p->amg, pcp, &p->mg[pcp]));
  }
  SolveWork *sw = solve_work ();
  ChebWork &cw = sw ? sw->cw[X.prec] : p->cw[X.prec];
  if (mixed) {
    if (X.prec == F64)
      return cheb_solve_mixed (*dz, *p->mg[pcp], p->lmin, p->lmax, p->nit, X,
                               Y, sw ? sw->cwm : p->cwm, yscale);
    return scaled (cheb_solve (*dz, p->mg[pcp], p->lmin, p->lmax, p->nit, X, Y, cw));
  }
  return scaled (cheb_solve (*dz, mode <= 3 ? p->mg[X.prec] : nullptr, p->lmin,
                             p->lmax, p->nit, X, Y, cw));
}

int
lgh::prior_solve_vec (lgh_prior_t *p, int trans, const PMat &X, PMat &Y)
{
  if (!p->is_mat) return cb_block (p, trans ? 1 : 0, X, Y);
  DevCSR *dz = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk (p->lazy_mu);
    LGH_TRY (prior_setup_device (p, X.prec));
    LGH_TRY (mat_device_csr (p->Z, X.prec, 1, &dz));
  }
  PcgInfo info;
  if (X.prec == F32) {
    // the PCG path keeps its FP64 accuracy: solve in FP64, convert
    PMat x64, y64;
    LGH_TRY (pmat_alloc (x64, X.nrows, X.ncols, X.W, F64));
    LGH_TRY (pmat_alloc (y64, X.nrows, X.ncols, X.W, F64));
    int rc = k_copy_cols (X, 0, x64, 0, X.ncols);
    if (!rc) rc = prior_solve_vec (p, trans, x64, y64);
    if (!rc) rc = k_copy_cols (y64, 0, Y, 0, X.ncols);
    pmat_free (x64);
    pmat_free (y64);
    return rc;
  }
  LGH_TRY (pcg_solve (*dz, p->amg_ok ? p->mg[F64] : nullptr, p->d_zdinv, X, Y,
                      p->o.ksp_rtol, 10000, &info));
  LGH_CHECK (info.converged, 10, "Z PCG solve did not converge in 10000 its");
  return 0;
}

int
lgh::prior_apply_Z (lgh_prior_t *p, int trans, const PMat &X, PMat &Y)
{
  if (!p->is_mat) return cb_block (p, trans ? 3 : 2, X, Y);
  LGH_TRY (prior_setup_device (p, X.prec));
  DevCSR *dz = nullptr;
  LGH_TRY (mat_device_csr (p->Z, X.prec, 1, &dz));
  return k_spmm (*dz, 1.0, X, 0.0, Y);
}

extern "C" int
lgh_prior_apply (lgh_prior_t *p, Vec x, Vec y)
{
  // y = Z M^{-1} Z^T x (glr_common_impl.h:236-244)
  if (!p->is_mat) {
    lgh_vec_fn at = p->cb.applyZt ? p->cb.applyZt : p->cb.applyZ;
    at (x, p->work, p->cb.ctx);
    VecPointwiseDivide (p->work, p->work, p->mass);
    p->cb.applyZ (p->work, y, p->cb.ctx);
    return 0;
  }
  PMat X, T;
  int rc = prior_setup_device (p, F64);
  if (!rc) rc = pmat_alloc (X, p->n, 1, 1, F64);
  if (!rc) rc = pmat_alloc (T, p->n, 1, 1, F64);
  if (!rc) rc = h2d (X.p, x->a, sizeof (double) * p->n);
  if (!rc) rc = prior_apply_Z (p, 1, X, T);
  if (!rc) rc = k_diag_scale (T, p->d_minv, T);
  if (!rc) rc = prior_apply_Z (p, 0, T, X);
  if (!rc) rc = d2h (y->a, X.p, sizeof (double) * p->n);
  pmat_free (X);
  pmat_free (T);
  if (rc) return lgh_petsc_error (PETSC_ERR_LIB, "lgh_prior_apply: %s",
                                  lgh::last_error ());
  return 0;
}

extern "C" int
lgh_prior_get_info (const lgh_prior_t *p, lgh_prior_info_t *info)
{
  memset (info, 0, sizeof (*info));
  info->is_mat = p->is_mat;
  info->blocked_mode = p->o.blocked_mode;
  info->nit = p->nit;
  info->lmin = p->lmin;
  info->lmax = p->lmax;
  info->levels = (int) p->amg.lev.size ();
  info->n_coarse = p->amg.lev.empty () ? 0 : p->amg.lev[0].A.n;
  info->bound_its = p->bound_its;
  info->cycle = p->amg.gamma;
  return 0;
}

extern "C" int
lgh_prior_solve_host (lgh_prior_t *p, int which, int prec, int m,
                      const double *X, int ldx, double *Y, int ldy)
{
  // prec: 0 FP64, 1 FP32, 2 FP64 vectors around the FP32 V-cycle,
  // 3 FP64 vectors around the BF16 V-cycle, 4 FP32 around the BF16 V-cycle
  const int n = p->n;
  if (prec < 0 || prec > 4)
    return lgh_petsc_error (PETSC_ERR_ARG_OUTOFRANGE,
                            "lgh_prior_solve_host: prec must be 0..4");
  const int vprec = (prec == 1 || prec == 4) ? F32 : F64;
  const int pcp = (prec == 2) ? F32 : (prec >= 3) ? BF16 : -1;
  std::vector<double> hx ((size_t) n * m);
  for (int j = 0; j < m; j++) memcpy (&hx[(size_t) j * n], X + (size_t) j * ldx, sizeof (double) * n);
  int W = 1;
  while (W < m && W < 64) W *= 2;
  if (prec >= 2) W = 64;
  PMat t, x, y;
  int rc = prior_setup_device (p, vprec);
  if (!rc) rc = pmat_alloc (t, n, m, 1, F64);
  if (!rc) rc = pmat_alloc (x, n, m, W, vprec);
  if (!rc) rc = pmat_alloc (y, n, m, W, vprec);
  if (!rc) rc = h2d (t.p, hx.data (), sizeof (double) * n * m);
  if (!rc) rc = k_copy_cols (t, 0, x, 0, m);
  if (!rc) rc = which == 0 ? prior_solve_block (p, 0, x, y, pcp) : prior_solve_vec (p, 0, x, y);
  if (!rc) rc = k_copy_cols (y, 0, t, 0, m);
  if (!rc) rc = d2h (hx.data (), t.p, sizeof (double) * n * m);
  pmat_free (t); pmat_free (x); pmat_free (y);
  if (rc) return lgh_petsc_error (PETSC_ERR_LIB, "lgh_prior_solve_host: %s",
                                  lgh::last_error ());
  for (int j = 0; j < m; j++) memcpy (Y + (size_t) j * ldy, &hx[(size_t) j * n], sizeof (double) * n);
  return 0;
}
