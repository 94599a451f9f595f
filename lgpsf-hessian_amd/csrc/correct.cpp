// correct.cpp — stage 2.5: deflation correction of a GLR against the true
// misfit Hessian (lgh_glr_correct_probes / lgh_glr_correct), following the
// algorithm of include/lgpsf_hessian/impl/correct_impl.h of the reference.
//
// Every n-row block lives in HBM as W = 64 FP64 panels; the n-row work is
// the engine's device kernels (Woodbury filters with U, MFMA Gram X^T Y and
// X C products, Z solves of the prior); only the small r x r / n2 x n2
// symmetric problems and the selection logic run on the host.  The true
// Hessian is the caller's host callback, applied column by column.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace lgh;

#define API_FAIL(what)                                                        \
  lgh_petsc_error (PETSC_ERR_LIB, "%s: %s", what, lgh::last_error ())

static inline int npan (int c) { return (c + 63) / 64; }

extern "C" lgh_glr_correct_opts_t
lgh_glr_correct_opts_default (void)
{
  // correct_impl.h:44-59
  lgh_glr_correct_opts_t o;
  o.c0 = 1.0;
  o.rank = -1;
  o.applies = 30;
  o.n_qc = 0;
  o.q_power = 0;
  o.clamp_eps = 0.05;
  o.d_keep_min = 0.;
  o.drop_tol = 1e-12;
  o.seed = 424242UL;
  o.verbose = 0;
  return o;
}

namespace {

// an n x m FP64 block of W = 64 panels, freed on scope exit
struct Blk {
  PMat p;
  int m = 0;
  Blk () = default;
  Blk (const Blk &) = delete;
  Blk &operator= (const Blk &) = delete;
  ~Blk () { pmat_free (p); }
  int alloc (int n, int cols)
  {
    pmat_free (p);
    m = cols;
    return pmat_alloc (p, n, std::max (cols, 1), 64, F64);
  }
};

// G (p x q, column-major, host) = X[:, :p]^T Y[:, :q]
int
xty (const Blk &X, const Blk &Y, std::vector<double> &G)
{
  const int p = X.m, q = Y.m, Px = npan (p), Py = npan (q);
  const int ldg = Px * 64;
  G.assign ((size_t) std::max (p * q, 1), 0.0);
  if (p == 0 || q == 0) return 0;
  double *d = (double *) dmalloc (sizeof (double) * (size_t) ldg * Py * 64);
  if (!d) return 5;
  std::vector<double> h ((size_t) ldg * Py * 64);
  int rc = k_gram (X.p, 0, Px, Y.p, 0, Py, d, ldg, 0);
  if (!rc) rc = d2h (h.data (), d, sizeof (double) * h.size ());
  dfree (d);
  LGH_TRY (rc);
  for (int j = 0; j < q; j++)
    for (int i = 0; i < p; i++) G[i + (size_t) j * p] = h[i + (size_t) j * ldg];
  return 0;
}

// Out[:, :q] = alpha A[:, :p] C + beta Out  (C: p x q column-major, ldc)
int
gemm (double alpha, const Blk &A, const double *C, int ldc, int q, double beta,
      Blk &Out)
{
  const int p = A.m, Pa = npan (p), Pq = npan (q), ld = Pa * 64;
  if (q == 0) return 0;
  std::vector<double> h ((size_t) ld * Pq * 64, 0.0);
  for (int j = 0; j < q; j++)
    for (int i = 0; i < p; i++) h[i + (size_t) j * ld] = C[i + (size_t) j * ldc];
  double *d = (double *) dmalloc (sizeof (double) * h.size ());
  if (!d) return 5;
  int rc = h2d (d, h.data (), sizeof (double) * h.size ());
  if (!rc && p > 0) rc = k_gemm_nn (A.p, 0, Pa, d, ld, alpha, beta, Out.p, 0, Pq);
  if (!rc && p == 0 && beta == 0.0) rc = dzero (Out.p.p, Out.p.bytes ());
  dfree (d);
  return rc;
}

// symmetric eigensolve of a small host matrix (column-major, n x n):
// eigenvectors in A, eigenvalues ascending in w
int
syevd_host (int n, std::vector<double> &A, std::vector<double> &w)
{
  w.assign (std::max (n, 1), 0.0);
  if (n == 0) return 0;
  double *d = (double *) dmalloc (sizeof (double) * ((size_t) n * n + n));
  if (!d) return 5;
  int info = 0;
  int rc = h2d (d, A.data (), sizeof (double) * (size_t) n * n);
  if (!rc) rc = sol_syevd (n, d, n, d + (size_t) n * n, &info);
  if (!rc) rc = d2h (A.data (), d, sizeof (double) * (size_t) n * n);
  if (!rc) rc = d2h (w.data (), d + (size_t) n * n, sizeof (double) * n);
  dfree (d);
  LGH_TRY (rc);
  LGH_CHECK (info == 0, 11, "lgh_glr_correct: symmetric eigensolve failed");
  return 0;
}

void
symmetrize (std::vector<double> &S, int r)
{
  for (int j = 0; j < r; j++)
    for (int i = 0; i < j; i++) {
      const double a = 0.5 * (S[i + (size_t) j * r] + S[j + (size_t) i * r]);
      S[i + (size_t) j * r] = a;
      S[j + (size_t) i * r] = a;
    }
}

// |values|-descending order by the swap sort of the reference (ties keep
// the first index found, as correct_impl.h:425-430)
std::vector<int>
abs_desc_order (const std::vector<double> &ev, int n)
{
  std::vector<int> o (n);
  for (int i = 0; i < n; i++) o[i] = i;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (std::fabs (ev[o[j]]) > std::fabs (ev[o[i]])) std::swap (o[i], o[j]);
  return o;
}

int
copy_cols (const Blk &A, int a0, Blk &B, int b0, int m)
{
  return k_copy_cols (A.p, a0, B.p, b0, m);
}

// Y = (U lam U^T + c0 I)^p X   (lgh_glr_pow_block on the current object)
int
pow_block (lgh_glr_t *g, double c0, double p, const Blk &X, Blk &Y)
{
  std::vector<double> f (std::max (g->kept, 1));
  for (int i = 0; i < g->kept; i++) f[i] = std::pow (g->lam[i] + c0, p);
  return glr_filter_panels (g, f.data (), std::pow (c0, p), X.p, X.m, Y.p);
}

// Y = F_true X = M^{1/2} Z^{-1} H_d Z^{-T} M^{1/2} X  (correct_impl.h:208-237)
int
ftrue_block (lgh_glr_t *g, lgh_hessian_fn hess, void *ctx, const Blk &X,
             Blk &Y, int *napplies)
{
  lgh_prior_t *p = g->prior;
  const int n = g->n, m = X.m;
  Blk W;
  PMat H;
  LGH_TRY (W.alloc (n, m));
  LGH_TRY (k_diag_scale (X.p, p->d_msqrt, W.p));
  LGH_TRY (prior_solve_block (p, 1, W.p, Y.p));
  int rc = pmat_alloc (H, n, m, 1, F64);
  std::vector<double> a ((size_t) n * m), b ((size_t) n * m);
  if (!rc) rc = k_copy_cols (Y.p, 0, H, 0, m);
  if (!rc) rc = d2h (a.data (), H.p, sizeof (double) * a.size ());
  if (!rc) {
    for (int j = 0; j < m; j++)
      hess (a.data () + (size_t) j * n, b.data () + (size_t) j * n, ctx);
    *napplies += m;
    rc = h2d (H.p, b.data (), sizeof (double) * b.size ());
  }
  if (!rc) rc = k_copy_cols (H, 0, W.p, 0, m);
  pmat_free (H);
  LGH_TRY (rc);
  LGH_TRY (prior_solve_block (p, 0, W.p, Y.p));
  return k_diag_scale (Y.p, p->d_msqrt, Y.p);
}

// E = E_w T = S^- ((F_true + c0 I)(S^- T)) - T   (correct_impl.h:239-250)
int
ew_block (lgh_glr_t *g, double c0, lgh_hessian_fn hess, void *ctx,
          const Blk &T, Blk &E, int *napplies)
{
  const int n = g->n, m = T.m;
  Blk S1, S2;
  LGH_TRY (S1.alloc (n, m));
  LGH_TRY (S2.alloc (n, m));
  LGH_TRY (E.alloc (n, m));
  LGH_TRY (pow_block (g, c0, -0.5, T, S1));
  LGH_TRY (ftrue_block (g, hess, ctx, S1, S2, napplies));
  LGH_TRY (k_axpby (c0, S1.p, 1.0, S2.p));
  LGH_TRY (pow_block (g, c0, -0.5, S2, E));
  return k_axpby (-1.0, T.p, 1.0, E.p);
}

// energy-ordered orthonormal basis of the columns of IMG: Gram
// eigendecomposition, drop energies <= drop_tol * max(top energy,
// scale_floor), columns in descending energy  (correct_impl.h:252-296)
int
energy_basis (const Blk &IMG, double drop_tol, double scale_floor, Blk &Q,
              int *dim)
{
  const int k = IMG.m;
  std::vector<double> G, ev;
  LGH_TRY (xty (IMG, IMG, G));
  LGH_TRY (syevd_host (k, G, ev));
  const double scale = std::max (k > 0 ? ev[k - 1] : 0.0, scale_floor);
  int nkeep = 0;
  for (int j = k - 1; j >= 0; j--)
    if (ev[j] > drop_tol * scale && ev[j] > 0.) nkeep++;
    else break;
  std::vector<double> C ((size_t) k * std::max (nkeep, 1), 0.0);
  for (int j = 0; j < nkeep; j++) {
    const int idx = k - 1 - j;
    const double s = 1.0 / std::sqrt (ev[idx]);
    for (int i = 0; i < k; i++) C[i + (size_t) j * k] = G[i + (size_t) idx * k] * s;
  }
  LGH_TRY (Q.alloc (IMG.p.nrows, nkeep));
  if (nkeep > 0) LGH_TRY (gemm (1.0, IMG, C.data (), k, nkeep, 0.0, Q));
  *dim = nkeep;
  return 0;
}

// whitened energy ratio of the current model on held-out pairs:
// sqrt(|U lam U^T Xw - Yw|_F^2 / |Yw|_F^2)   (correct_impl.h:298-321)
int
qc_ratio (lgh_glr_t *g, const Blk &Xw, const Blk &Yw, double *qc)
{
  Blk R;
  LGH_TRY (R.alloc (g->n, Xw.m));
  std::vector<double> f (std::max (g->kept, 1));
  for (int i = 0; i < g->kept; i++) f[i] = g->lam[i];
  LGH_TRY (glr_filter_panels (g, f.data (), 0.0, Xw.p, Xw.m, R.p));
  LGH_TRY (k_axpby (-1.0, Yw.p, 1.0, R.p));
  std::vector<double> Gr, Gy;
  LGH_TRY (xty (R, R, Gr));
  LGH_TRY (xty (Yw, Yw, Gy));
  double num = 0., den = 0.;
  for (int j = 0; j < Xw.m; j++) {
    num += Gr[j + (size_t) j * Xw.m];
    den += Gy[j + (size_t) j * Xw.m];
  }
  *qc = den > 0. ? std::sqrt (num / den) : 0.;
  return 0;
}

// fold Vhat diag(d) Vhat^T into (U, lam): orthogonalize Vhat against U,
// eigendecompose the combined coefficient matrix on [U | Qv], keep by
// |lam'| with the object's truncation rules, SIGNED (correct_impl.h:323-488)
int
graft (lgh_glr_t *g, const Blk &Vhat, const std::vector<double> &d,
       const lgh_glr_correct_opts_t &o, lgh_glr_correct_report_t *rep)
{
  // an implicit-U object gets its U = Q V formed first
  // (correct_impl.h:338-345); the sketch state is dropped below anyway
  LGH_TRY (glr_materialize_U (g));
  const int n = g->n, r = Vhat.m, k0 = g->kept;
  std::vector<double> Gv;
  LGH_TRY (xty (Vhat, Vhat, Gv));
  double scale_floor = 0.;
  for (int i = 0; i < r; i++) scale_floor += Gv[i + (size_t) i * r];
  scale_floor /= std::max (r, 1);
  // Vperp = (I - U U^T) Vhat
  Blk Vperp, Ub;
  LGH_TRY (Vperp.alloc (n, r));
  LGH_TRY (copy_cols (Vhat, 0, Vperp, 0, r));
  std::vector<double> Cuv;
  if (k0 > 0) {
    Ub.p = g->U;            // borrowed: released before Ub goes out of scope
    Ub.m = k0;
    int rc = xty (Ub, Vhat, Cuv);
    if (!rc) rc = gemm (-1.0, Ub, Cuv.data (), k0, r, 1.0, Vperp);
    Ub.p = PMat ();
    LGH_TRY (rc);
  }
  Blk Qv;
  int r2 = 0;
  LGH_TRY (energy_basis (Vperp, o.drop_tol, scale_floor, Qv, &r2));
  std::vector<double> Cqv;
  if (r2 > 0) LGH_TRY (xty (Qv, Vhat, Cqv));
  // T2 = diag(lam, 0) + CB diag(d) CB^T,  CB = [Cuv; Cqv]  (n2 x r)
  const int n2 = k0 + r2;
  std::vector<double> T2 ((size_t) std::max (n2 * n2, 1), 0.0), CB ((size_t) n2 * r);
  for (int i = 0; i < k0; i++) T2[i + (size_t) i * n2] = g->lam[i];
  for (int jc = 0; jc < r; jc++) {
    for (int i = 0; i < k0; i++) CB[i + (size_t) jc * n2] = Cuv[i + (size_t) jc * k0];
    for (int i = 0; i < r2; i++) CB[k0 + i + (size_t) jc * n2] = Cqv[i + (size_t) jc * r2];
  }
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n2; i++) {
      double s = 0.;
      for (int jc = 0; jc < r; jc++)
        s += CB[i + (size_t) jc * n2] * d[jc] * CB[j + (size_t) jc * n2];
      T2[i + (size_t) j * n2] += s;
    }
  symmetrize (T2, n2);
  std::vector<double> ev2;
  LGH_TRY (syevd_host (n2, T2, ev2));
  // |lam'|-descending, the object's truncation rules, signed
  const std::vector<int> ord = abs_desc_order (ev2, n2);
  const double absmax = n2 > 0 ? std::fabs (ev2[ord[0]]) : 0.;
  double cut = 0.;
  if (g->opts.trunc_abs > 0.) cut = std::max (cut, g->opts.trunc_abs);
  if (g->opts.trunc_rel > 0.) cut = std::max (cut, g->opts.trunc_rel * absmax);
  int keptn = 0;
  while (keptn < n2 && std::fabs (ev2[ord[keptn]]) > cut) keptn++;
  std::vector<double> lam_new (keptn);
  double lmin = 0.;
  for (int i = 0; i < keptn; i++) {
    lam_new[i] = ev2[ord[i]];
    lmin = std::min (lmin, lam_new[i]);
  }
  // U' = U W[0:k0] + Qv W[k0:n2]
  std::vector<double> Wsel ((size_t) std::max (n2 * keptn, 1), 0.0);
  for (int j = 0; j < keptn; j++)
    for (int i = 0; i < n2; i++) Wsel[i + (size_t) j * n2] = T2[i + (size_t) ord[j] * n2];
  Blk Un;
  LGH_TRY (Un.alloc (n, keptn));
  if (k0 > 0 && keptn > 0) {
    Ub.p = g->U;
    Ub.m = k0;
    int rc = gemm (1.0, Ub, Wsel.data (), n2, keptn, 0.0, Un);
    Ub.p = PMat ();
    LGH_TRY (rc);
  }
  if (r2 > 0 && keptn > 0)
    LGH_TRY (gemm (1.0, Qv, Wsel.data () + k0, n2, keptn, k0 > 0 ? 1.0 : 0.0, Un));
  pmat_free (g->U);
  g->U = Un.p;
  Un.p = PMat ();
  g->lam = lam_new;
  g->kept = keptn;
  g->corrected = 1;
  g->floor = std::max (0., -lmin);
  rep->floor = g->floor;
  // the sketch state no longer describes the object (extend is refused)
  pmat_free (g->Q);
  g->Araw.clear ();
  return 0;
}

// Rayleigh-price the orthonormal directions QB with images EB = E_w QB,
// clamp, keep by |d|, Vhat = S^+ (QB Vs), graft  (correct_impl.h:490-572)
int
price_and_graft (lgh_glr_t *g, const lgh_glr_correct_opts_t &o,
                 lgh_glr_correct_report_t *rep, const Blk &QB, const Blk &EB)
{
  const int r = QB.m, n = g->n;
  const double t0 = MPI_Wtime ();
  std::vector<double> S, ev;
  LGH_TRY (xty (QB, EB, S));
  symmetrize (S, r);
  LGH_TRY (syevd_host (r, S, ev));
  const std::vector<int> ord = abs_desc_order (ev, r);
  int rkeep = (o.rank < 0) ? r : std::min (r, o.rank);
  while (rkeep > 0 && std::fabs (ev[ord[rkeep - 1]]) < o.d_keep_min) rkeep--;
  if (rkeep == 0) {
    // nothing clears the threshold: the object is left as it was
    rep->kept = 0;
    rep->clamped = 0;
    rep->d_min = r > 0 ? ev[0] : 0.;
    rep->d_max = r > 0 ? ev[r - 1] : 0.;
    rep->floor = 0.;
    return 0;
  }
  std::vector<double> dsel (rkeep), Csel ((size_t) r * rkeep);
  rep->clamped = 0;
  for (int j = 0; j < rkeep; j++) {
    double dj = ev[ord[j]];
    if (dj < -1. + o.clamp_eps) {
      dj = -1. + o.clamp_eps;
      rep->clamped++;
    }
    dsel[j] = dj;
    rep->d_min = j == 0 ? dj : std::min (rep->d_min, dj);
    rep->d_max = j == 0 ? dj : std::max (rep->d_max, dj);
    for (int i = 0; i < r; i++) Csel[i + (size_t) j * r] = S[i + (size_t) ord[j] * r];
  }
  Blk Vw, Vhat;
  LGH_TRY (Vw.alloc (n, rkeep));
  LGH_TRY (Vhat.alloc (n, rkeep));
  LGH_TRY (gemm (1.0, QB, Csel.data (), r, rkeep, 0.0, Vw));
  LGH_TRY (pow_block (g, o.c0, 0.5, Vw, Vhat));
  LGH_TRY (graft (g, Vhat, dsel, o, rep));
  rep->kept = rkeep;
  rep->t_dense += MPI_Wtime () - t0;
  return 0;
}

// X = M^{-1/2} Z^T V and Yw = M^{1/2} Z^{-1} HV on m host columns
int
transform_probes (lgh_glr_t *g, const double *V, const double *HV, int m,
                  Blk &X, Blk &Yw)
{
  lgh_prior_t *p = g->prior;
  const int n = g->n;
  PMat H;
  Blk T;
  LGH_TRY (X.alloc (n, m));
  LGH_TRY (Yw.alloc (n, m));
  LGH_TRY (T.alloc (n, m));
  int rc = pmat_alloc (H, n, m, 1, F64);
  if (!rc) rc = h2d (H.p, V, sizeof (double) * (size_t) n * m);
  if (!rc) rc = k_copy_cols (H, 0, T.p, 0, m);
  if (!rc) rc = prior_apply_Z (p, 1, T.p, X.p);
  if (!rc) rc = k_diag_scale (X.p, p->d_minvsqrt, X.p);
  if (!rc) rc = h2d (H.p, HV, sizeof (double) * (size_t) n * m);
  if (!rc) rc = k_copy_cols (H, 0, T.p, 0, m);
  if (!rc) rc = prior_solve_block (p, 0, T.p, Yw.p);
  if (!rc) rc = k_diag_scale (Yw.p, p->d_msqrt, Yw.p);
  pmat_free (H);
  return rc;
}

int
correct_probes (lgh_glr_t *g, int k, const double *V, const double *HV,
                lgh_hessian_fn hess, void *ctx, const lgh_glr_correct_opts_t &o,
                lgh_glr_correct_report_t *rep)
{
  const int n = g->n, kf = k - o.n_qc;
  LGH_TRY (prior_setup_device (g->prior, F64));
  double t0 = MPI_Wtime ();
  Blk X, Yw, Xq, Yqw;
  LGH_TRY (transform_probes (g, V, HV, kf, X, Yw));
  rep->t_operator += MPI_Wtime () - t0;
  if (o.n_qc > 0) {
    LGH_TRY (transform_probes (g, V + (size_t) kf * n, HV + (size_t) kf * n,
                               o.n_qc, Xq, Yqw));
    LGH_TRY (qc_ratio (g, Xq, Yqw, &rep->qc_before));
  }
  // residual Rw = Yw - U lam U^T X and its whitened image: the error basis,
  // free of true applies
  t0 = MPI_Wtime ();
  Blk Rw, IMG, QB0;
  LGH_TRY (Rw.alloc (n, kf));
  LGH_TRY (IMG.alloc (n, kf));
  {
    std::vector<double> f (std::max (g->kept, 1));
    for (int i = 0; i < g->kept; i++) f[i] = g->lam[i];
    LGH_TRY (glr_filter_panels (g, f.data (), 0.0, X.p, kf, Rw.p));
  }
  LGH_TRY (k_axpby (1.0, Yw.p, -1.0, Rw.p));
  LGH_TRY (pow_block (g, o.c0, -0.5, Rw, IMG));
  int bdim = 0;
  LGH_TRY (energy_basis (IMG, o.drop_tol, 0., QB0, &bdim));
  rep->basis = bdim;
  rep->t_dense += MPI_Wtime () - t0;
  if (bdim == 0) return -PETSC_ERR_CONV_FAILED;

  // price the top m1 basis directions (m1 true applies), then one power
  // step past the basis when the budget allows (correct_impl.h:700-761)
  int napplies = 0;
  const int m1 = std::min (o.applies, bdim);
  Blk Q1, E1, QB, EB;
  LGH_TRY (Q1.alloc (n, m1));
  LGH_TRY (copy_cols (QB0, 0, Q1, 0, m1));
  t0 = MPI_Wtime ();
  LGH_TRY (ew_block (g, o.c0, hess, ctx, Q1, E1, &napplies));
  rep->t_operator += MPI_Wtime () - t0;
  bool concat = false;
  if (o.applies > m1 && m1 > 0) {
    Blk P, Q2b;
    LGH_TRY (P.alloc (n, m1));
    LGH_TRY (copy_cols (E1, 0, P, 0, m1));
    std::vector<double> C, Ge;
    LGH_TRY (xty (Q1, E1, C));
    LGH_TRY (gemm (-1.0, Q1, C.data (), m1, m1, 1.0, P));
    LGH_TRY (xty (E1, E1, Ge));
    double sf = 0.;
    for (int j = 0; j < m1; j++) sf += Ge[j + (size_t) j * m1];
    sf /= std::max (m1, 1);
    int b2 = 0;
    LGH_TRY (energy_basis (P, o.drop_tol, sf, Q2b, &b2));
    const int m2 = std::min (o.applies - m1, b2);
    if (m2 > 0) {
      Blk Q2, E2;
      LGH_TRY (Q2.alloc (n, m2));
      LGH_TRY (copy_cols (Q2b, 0, Q2, 0, m2));
      t0 = MPI_Wtime ();
      LGH_TRY (ew_block (g, o.c0, hess, ctx, Q2, E2, &napplies));
      rep->t_operator += MPI_Wtime () - t0;
      LGH_TRY (QB.alloc (n, m1 + m2));
      LGH_TRY (EB.alloc (n, m1 + m2));
      LGH_TRY (copy_cols (Q1, 0, QB, 0, m1));
      LGH_TRY (copy_cols (Q2, 0, QB, m1, m2));
      LGH_TRY (copy_cols (E1, 0, EB, 0, m1));
      LGH_TRY (copy_cols (E2, 0, EB, m1, m2));
      concat = true;
    }
  }
  LGH_TRY (price_and_graft (g, o, rep, concat ? QB : Q1, concat ? EB : E1));
  rep->hd_applies = napplies;
  if (o.n_qc > 0) LGH_TRY (qc_ratio (g, Xq, Yqw, &rep->qc_after));
  return 0;
}

int
correct_fresh (lgh_glr_t *g, lgh_hessian_fn hess, void *ctx,
               const lgh_glr_correct_opts_t &o, lgh_glr_correct_report_t *rep)
{
  const int n = g->n;
  const int passes = 2 + std::max (o.q_power, 0);
  const int mr = std::max (o.applies / passes, 1);
  LGH_TRY (prior_setup_device (g->prior, F64));
  int napplies = 0;
  // hashed Gaussian test block in the whitened frame (correct_impl.h:821-832)
  Blk T, IMG;
  LGH_TRY (T.alloc (n, mr));
  LGH_TRY (k_randn (T.p, o.seed, 0, 0));     // padding columns left zero
  double t0 = MPI_Wtime ();
  LGH_TRY (ew_block (g, o.c0, hess, ctx, T, IMG, &napplies));
  rep->t_operator += MPI_Wtime () - t0;
  for (int q = 0; q < o.q_power; q++) {
    Blk Tn;
    int bd = 0;
    LGH_TRY (energy_basis (IMG, o.drop_tol, 0., Tn, &bd));
    if (bd == 0) break;
    t0 = MPI_Wtime ();
    LGH_TRY (ew_block (g, o.c0, hess, ctx, Tn, IMG, &napplies));
    rep->t_operator += MPI_Wtime () - t0;
  }
  Blk QB0, QB, EB;
  int bdim = 0;
  LGH_TRY (energy_basis (IMG, o.drop_tol, 0., QB0, &bdim));
  rep->basis = bdim;
  if (bdim == 0) return -PETSC_ERR_CONV_FAILED;
  const int m1 = std::min (std::max (o.applies - napplies, 1), bdim);
  LGH_TRY (QB.alloc (n, m1));
  LGH_TRY (copy_cols (QB0, 0, QB, 0, m1));
  t0 = MPI_Wtime ();
  LGH_TRY (ew_block (g, o.c0, hess, ctx, QB, EB, &napplies));
  rep->t_operator += MPI_Wtime () - t0;
  LGH_TRY (price_and_graft (g, o, rep, QB, EB));
  rep->hd_applies = napplies;
  return 0;
}

PetscErrorCode
finish (int rc, const char *name, lgh_glr_correct_report_t &rep,
        lgh_glr_correct_report_t *report, const lgh_glr_correct_opts_t &o)
{
  if (rc == -PETSC_ERR_CONV_FAILED)
    return lgh_petsc_error (PETSC_ERR_CONV_FAILED,
                            "%s: the error probes span nothing — the "
                            "approximation is already exact on them or the "
                            "probe data is inconsistent", name);
  if (rc) return API_FAIL (name);
  if (o.verbose)
    PetscPrintf (PETSC_COMM_SELF, "[lgh correct] basis %d, priced %d applies, "
                 "kept %d (clamped %d), d in [%.3g, %.3g], floor %.3g, qc %.3g "
                 "-> %.3g\n", rep.basis, rep.hd_applies, rep.kept, rep.clamped,
                 rep.d_min, rep.d_max, rep.floor, rep.qc_before, rep.qc_after);
  if (report) *report = rep;
  return 0;
}

}  // namespace

extern "C" int
lgh_glr_correct_probes (lgh_glr_t *glr, int k, const double *V,
                        const double *HV, lgh_hessian_fn hessian_apply,
                        void *ctx, const lgh_glr_correct_opts_t *opts,
                        lgh_glr_correct_report_t *report)
{
  const lgh_glr_correct_opts_t o = opts ? *opts : lgh_glr_correct_opts_default ();
  PetscCheck (glr != NULL && V != NULL && HV != NULL && hessian_apply != NULL,
              0, PETSC_ERR_ARG_NULL, "lgh_glr_correct_probes: NULL argument");
  // correct_impl.h:610-618
  PetscCheck (!glr->corrected, 0, PETSC_ERR_ORDER,
              "lgh_glr_correct_probes: the object is already corrected");
  PetscCheck (k - o.n_qc > 0, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_correct_probes: need k - n_qc > 0 probes (k=%d, n_qc=%d)",
              k, o.n_qc);
  PetscCheck (o.applies > 0, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_correct_probes: applies budget must be positive");
  PetscCheck (o.c0 > 0., 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_correct_probes: c0 must be positive");
  lgh_glr_correct_report_t rep;
  memset (&rep, 0, sizeof (rep));
  rep.qc_before = -1.;
  rep.qc_after = -1.;
  const int rc = correct_probes (glr, k, V, HV, hessian_apply, ctx, o, &rep);
  return finish (rc, "lgh_glr_correct_probes", rep, report, o);
}

extern "C" int
lgh_glr_correct (lgh_glr_t *glr, lgh_hessian_fn hessian_apply, void *ctx,
                 const lgh_glr_correct_opts_t *opts,
                 lgh_glr_correct_report_t *report)
{
  const lgh_glr_correct_opts_t o = opts ? *opts : lgh_glr_correct_opts_default ();
  PetscCheck (glr != NULL && hessian_apply != NULL, 0, PETSC_ERR_ARG_NULL,
              "lgh_glr_correct: NULL argument");
  // correct_impl.h:805-813
  PetscCheck (!glr->corrected, 0, PETSC_ERR_ORDER,
              "lgh_glr_correct: the object is already corrected");
  PetscCheck (o.applies >= 2, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_correct: the fresh path needs applies >= 2 (range pass "
              "+ Rayleigh pass)");
  PetscCheck (o.c0 > 0., 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_correct: c0 must be positive");
  lgh_glr_correct_report_t rep;
  memset (&rep, 0, sizeof (rep));
  rep.qc_before = -1.;
  rep.qc_after = -1.;
  const int rc = correct_fresh (glr, hessian_apply, ctx, o, &rep);
  return finish (rc, "lgh_glr_correct", rep, report, o);
}
