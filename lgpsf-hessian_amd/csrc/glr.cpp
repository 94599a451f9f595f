// glr.cpp — lgh_glr_compute / _extend / downstream operations on the GPU.
//
// Host code keeps the reference control flow (glr_common_impl.h:409-1104;
// incremental extension of glr_scalapack_impl.h:689-915); every O(n)
// array lives in HBM as a W = 64 panel matrix and only ell x ell blocks,
// kept-sized coefficient vectors and scalars cross back to the host.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace lgh;

#define API_FAIL(what)                                                        \
  lgh_petsc_error (PETSC_ERR_LIB, "%s: %s", what, lgh::last_error ())

extern "C" lgh_glr_opts_t
lgh_glr_opts_default (void)
{
  // glr_common_impl.h:295-312
  lgh_glr_opts_t o;
  o.ell = 200;
  o.q_power = 1;
  o.trunc_abs = 0.;
  o.trunc_rel = 1e-3;
  o.treat = LGH_TREAT_FLIP;
  o.seed = 271828UL;
  o.backend = LGH_GLR_DEFAULT;
  o.nb = 64;
  o.panel = 64;
  o.grid2d = 0;
  o.check = 1;
  return o;
}

extern "C" lgh_glr_xopts_t
lgh_glr_xopts_default (void)
{
  lgh_glr_xopts_t x;
  x.prec_sketch = LGH_PREC_F64;
  x.prec_precond = LGH_PREC_F64;
  x.prec_check = LGH_PREC_F64;
  x.prec_sketch_precond = LGH_PREC_F64;
  x.solve_blocked_min = 0;
  x.sweep_panels = 0;
  return x;
}

extern "C" lgh_glr_xopts_t
lgh_glr_xopts_mixed (void)
{
  lgh_glr_xopts_t x;
  x.prec_sketch = LGH_PREC_F64;
  x.prec_precond = LGH_PREC_BF16;
  x.prec_check = LGH_PREC_F32;
  x.prec_sketch_precond = LGH_PREC_BF16;
  x.solve_blocked_min = 1;
  x.sweep_panels = 0;
  return x;
}

static inline int to_prec (int p)
{
  return p == LGH_PREC_F32 ? F32 : p == LGH_PREC_BF16 ? BF16 : F64;
}
// V-cycle precision of a sweep whose vectors are in vp: the policy's
// V-cycle precision for such sweeps when it is lower than vp, else vp
static inline int pc_for (const lgh_glr_t *g, int vp)
{
  auto rank = [] (int q) { return q == F64 ? 2 : q == F32 ? 1 : 0; };
  const int pp = to_prec (vp == F64 ? g->x.prec_precond : g->x.prec_sketch_precond);
  return rank (pp) < rank (vp) ? pp : vp;
}

static inline int npan (int c) { return (c + 63) / 64; }

// ---- F sweep (glr_common_impl.h:316-336) -----------------------------------
// Y = Msqrt Z^{-1} B Z^{-T} Msqrt X ; Wk is scratch of X's shape
static int
apply_F (lgh_glr_t *g, const PMat &X, PMat &Y, PMat &Wk, int pcp = -1)
{
  lgh_prior_t *p = g->prior;
  DevCSR *dB = nullptr;
  LGH_TRY (prior_setup_device (p, X.prec));
  LGH_TRY (mat_device_csr (g->B, X.prec, 0, &dB));
  LGH_TRY (k_diag_scale (X, p->d_msqrt, Wk));
  LGH_TRY (prior_solve_block (p, 1, Wk, Y, pcp));
  prof_set_b_tag (1);
  int rc = k_spmm (*dB, 1.0, Y, 0.0, Wk);
  prof_set_b_tag (0);
  LGH_TRY (rc);
  // Y = M^{1/2} Z^{-1} Wk (the scaling from the solve's last step)
  LGH_TRY (prior_solve_block (p, 0, Wk, Y, pcp, p->d_msqrt));
  return 0;
}

// ---- CholQR2 (glr_common_impl.h:340-407) -----------------------------------
// X <- X L^{-T}, twice; Gram on the device (MFMA, fixed-order reduction),
// Cholesky and the triangular inverse in FP64 (rocSOLVER), the triangular
// update as an MFMA product with the explicit L^{-T}.  `passes` = 2 is the
// reference's CholQR2.
static int
cholqr (PMat &X, int ncols, int passes, int *jittered)
{
  const int P = X.npanels (), ld = P * 64;
  double *G = (double *) dmalloc (sizeof (double) * (size_t) ld * ld);
  double *C = (double *) dmalloc (sizeof (double) * (size_t) ld * ld);
  double *sh = (double *) dmalloc (sizeof (double) * 2);
  PMat T;
  int rc = (G && C && sh) ? 0 : 5;
  if (!rc) rc = pmat_alloc (T, X.nrows, X.ncols, 64, X.prec);
  for (int pass = 0; pass < passes && !rc; pass++) {
    int info = 0;
    for (int attempt = 0; attempt < 2 && !rc; attempt++) {
      rc = k_gram (X, 0, P, X, 0, P, G, ld, 1);
      if (!rc && attempt == 1) {
        // re-formed Gram + 1e-12 trace/ell on the diagonal (:389-395)
        double tr = 0.0;
        rc = k_trace (G, ncols, ld, sh);
        if (!rc) rc = d2h (&tr, sh, sizeof (double));
        double s = 1e-12 * tr / ncols;
        if (!rc) rc = h2d (sh + 1, &s, sizeof (double));
        if (!rc) rc = k_add_diag (G, ncols, ld, sh + 1);
        if (jittered) *jittered = 1;
      }
      if (!rc) rc = sol_potrf (0, ncols, G, ld, &info);
      if (!rc && info == 0) break;
    }
    if (!rc && info != 0) {
      set_error ("CholQR2: potrf failed twice");
      rc = 11;
    }
    if (!rc) rc = sol_trtri (0, ncols, G, ld, &info);
    if (!rc && info != 0) { set_error ("CholQR2: trtri failed"); rc = 11; }
    if (!rc) rc = k_lower_t_pad (G, ncols, ld, C, ld, ld);
    if (!rc) rc = k_gemm_nn (X, 0, P, C, ld, 1.0, 0.0, T, 0, P, 1);
    if (!rc) std::swap (X.p, T.p);
  }
  pmat_free (T);
  dfree (G); dfree (C); dfree (sh);
  return rc;
}

// panel CholQR2 of glr_scalapack_impl.h:110-148 on one 64-column panel q of
// X holding bw valid columns (upper factor, jitter 1e-12 max(diag, 1))
static int
panel_cholqr (PMat &X, int q, int bw)
{
  double *G = (double *) dmalloc (sizeof (double) * 64 * 64);
  double *C = (double *) dmalloc (sizeof (double) * 64 * 64);
  int rc = (G && C) ? 0 : 5;
  for (int pass = 0; pass < 2 && !rc; pass++) {
    int info = 0;
    rc = k_gram (X, q, 1, X, q, 1, G, 64, 1);
    if (!rc) rc = sol_potrf (1, bw, G, 64, &info);
    if (!rc && info != 0) {
      std::vector<double> hg (64 * 64);
      rc = k_gram (X, q, 1, X, q, 1, G, 64, 1);
      if (!rc) rc = d2h (hg.data (), G, sizeof (double) * 64 * 64);
      double dmax = 0.0;
      for (int i = 0; i < bw; i++) dmax = std::max (dmax, hg[i + i * 64]);
      for (int i = 0; i < bw; i++) hg[i + i * 64] += 1.e-12 * std::max (dmax, 1.);
      if (!rc) rc = h2d (G, hg.data (), sizeof (double) * 64 * 64);
      if (!rc) rc = sol_potrf (1, bw, G, 64, &info);
      if (!rc && info != 0) { set_error ("panel CholQR: potrf failed twice"); rc = 11; }
    }
    if (!rc) rc = sol_trtri (1, bw, G, 64, &info);
    if (!rc) rc = k_upper_pad (G, bw, 64, C, 64, 64);
    if (!rc) rc = k_gemm_nn (X, q, 1, C, 64, 1.0, 0.0, X, q, 1);   // in place
  }
  dfree (G); dfree (C);
  return rc;
}

// BCGS2 (glr_scalapack_impl.h:153-195) of the k_new columns of Yn against
// the first `ncols_old` columns of Qold (may be null) and each other, in
// 64-column panels
static int
bcgs2 (const PMat *Qold, int ncols_old, PMat &Yn, int k_new)
{
  const int Po = Qold ? npan (ncols_old) : 0, Pn = npan (k_new);
  const int ldo = std::max (Po, 1) * 64, ldn = Pn * 64;
  double *C1 = (double *) dmalloc (sizeof (double) * (size_t) ldo * 64);
  double *C2 = (double *) dmalloc (sizeof (double) * (size_t) ldn * 64);
  int rc = (C1 && C2) ? 0 : 5;
  for (int t = 0; t < Pn && !rc; t++) {
    const int bw = std::min (64, k_new - 64 * t);
    for (int sweep = 0; sweep < 2 && !rc; sweep++) {
      if (Po > 0) {
        rc = k_gram (*Qold, 0, Po, Yn, t, 1, C1, ldo, 0);
        if (!rc) rc = k_gemm_nn (*Qold, 0, Po, C1, ldo, -1.0, 1.0, Yn, t, 1);
      }
      if (!rc && t > 0) {
        rc = k_gram (Yn, 0, t, Yn, t, 1, C2, ldn, 0);
        if (!rc) rc = k_gemm_nn (Yn, 0, t, C2, ldn, -1.0, 1.0, Yn, t, 1);
      }
      if (!rc) rc = panel_cholqr (Yn, t, bw);
    }
  }
  dfree (C1); dfree (C2);
  return rc;
}

// selection + treatment (glr_common_impl.h:517-559) from ascending evals
static void
select_modes (lgh_glr_t *g, int ell, const std::vector<double> &ev,
              lgh_glr_report_t *rep, std::vector<int> &order)
{
  order.resize (ell);
  for (int i = 0; i < ell; i++) order[i] = i;
  for (int i = 0; i < ell; i++)
    for (int j = i + 1; j < ell; j++)
      if (std::fabs (ev[order[j]]) > std::fabs (ev[order[i]])) std::swap (order[i], order[j]);
  const double absmax = std::fabs (ev[order[0]]);
  double cut = 0.;
  if (g->opts.trunc_abs > 0.) cut = std::max (cut, g->opts.trunc_abs);
  if (g->opts.trunc_rel > 0.) cut = std::max (cut, g->opts.trunc_rel * absmax);
  int kept = 0;
  while (kept < ell && std::fabs (ev[order[kept]]) > cut) kept++;
  rep->kept = kept;
  rep->lam_abs_max = absmax;
  rep->next_abs = (kept < ell) ? std::fabs (ev[order[kept]]) : 0.;
  rep->lam_raw_max = ev[ell - 1];
  rep->lam_raw_min = ev[0];
  rep->n_negative_raw = 0;
  for (int i = 0; i < ell; i++) if (ev[i] < 0.) rep->n_negative_raw++;
  g->lam.assign ((size_t) std::max (kept, 1), 0.0);
  for (int j = 0; j < kept; j++) {
    const double l = ev[order[j]];
    switch (g->opts.treat) {
    case LGH_TREAT_FLIP: g->lam[j] = std::fabs (l); break;
    case LGH_TREAT_RELU: g->lam[j] = std::max (l, 0.); break;
    default: g->lam[j] = l; break;
    }
  }
  g->kept = kept;
}

// V (ell x kept) to the device, zero-padded to 64-multiples
static int
upload_V (lgh_glr_t *g, int ell)
{
  const int k = g->kept, ldl = npan (ell) * 64, ldk = npan (std::max (k, 1)) * 64;
  dfree (g->dV);
  g->dV = (double *) dmalloc (sizeof (double) * (size_t) ldl * ldk);
  if (!g->dV) return 5;
  std::vector<double> h ((size_t) ldl * ldk, 0.0);
  for (int j = 0; j < k; j++)
    for (int i = 0; i < ell; i++) h[i + (size_t) j * ldl] = g->V[i + (size_t) j * ell];
  return h2d (g->dV, h.data (), sizeof (double) * h.size ());
}

// symmetrized eig of the raw projection A (ell x ell, host), selection,
// U = [Qa | Qb] C.  Qb (may be null) holds columns ell_a..ell-1.  An
// implicit object keeps V = C instead of forming U.  xcheck: the
// implicit-U backend's check (glr_scalapack_impl.h:633-670): the
// eigenvalues again by a second method (values-only QR iteration), and
// rep->resid_max = max |dlam| / max |lam| between the two
static int
rayleigh_ritz (lgh_glr_t *g, int ell, const std::vector<double> &Araw,
               const PMat &Qa, int ell_a, const PMat *Qb,
               lgh_glr_report_t *rep, std::vector<double> &evals,
               std::vector<int> &order, double *t_dense, int xcheck = 0)
{
  std::vector<double> T ((size_t) ell * ell);
  for (int j = 0; j < ell; j++)
    for (int i = 0; i < ell; i++) T[i + (size_t) j * ell] = Araw[i + (size_t) j * ell];
  for (int j = 0; j < ell; j++)          // glr_common_impl.h:490-494
    for (int i = 0; i < j; i++) {
      const double avg = 0.5 * (T[i + (size_t) j * ell] + T[j + (size_t) i * ell]);
      T[i + (size_t) j * ell] = avg;
      T[j + (size_t) i * ell] = avg;
    }
  const double t0 = MPI_Wtime ();
  double *dT = (double *) dmalloc (sizeof (double) * (size_t) ell * ell);
  double *dw = (double *) dmalloc (sizeof (double) * (size_t) ell);
  if (!dT || !dw) { dfree (dT); dfree (dw); return 5; }
  int info = 0;
  std::vector<double> evr;
  int rc = 0;
  if (xcheck) {                          // values-only pass on a copy
    evr.resize (ell);
    rc = h2d (dT, T.data (), sizeof (double) * (size_t) ell * ell);
    if (!rc) rc = sol_syev_vals (ell, dT, ell, dw, &info);
    if (!rc && info != 0) { set_error ("syev failed"); rc = 12; }
    if (!rc) rc = d2h (evr.data (), dw, sizeof (double) * ell);
  }
  if (!rc) rc = h2d (dT, T.data (), sizeof (double) * (size_t) ell * ell);
  if (!rc) rc = sol_syevd (ell, dT, ell, dw, &info);
  if (!rc && info != 0) { set_error ("syevd failed"); rc = 12; }
  evals.resize (ell);
  if (!rc) rc = d2h (evals.data (), dw, sizeof (double) * ell);
  if (!rc) rc = d2h (T.data (), dT, sizeof (double) * (size_t) ell * ell);
  dfree (dT); dfree (dw);
  *t_dense += MPI_Wtime () - t0;
  LGH_TRY (rc);
  if (xcheck) {
    double dmax = 0., amax = 0.;
    for (int i = 0; i < ell; i++) {
      dmax = std::max (dmax, std::fabs (evr[i] - evals[i]));
      amax = std::max (amax, std::fabs (evr[i]));
    }
    rep->resid_max = dmax / std::max (amax, 1e-300);
  }
  select_modes (g, ell, evals, rep, order);
  const int k = g->kept;
  if (g->implicit) {
    pmat_free (g->U);
    g->V.assign ((size_t) ell * k, 0.0);
    for (int j = 0; j < k; j++)
      for (int i = 0; i < ell; i++)
        g->V[i + (size_t) j * ell] = T[i + (size_t) order[j] * ell];
    return upload_V (g, ell);
  }
  pmat_free (g->U);
  LGH_TRY (pmat_alloc (g->U, g->n, std::max (k, 1), 64, F64));
  if (k == 0) return 0;
  const int Pk = npan (k), Pa = npan (ell_a);
  const int lda = Pa * 64, ldk = Pk * 64;
  std::vector<double> Ca ((size_t) lda * ldk, 0.0), Cb;
  for (int j = 0; j < k; j++)
    for (int i = 0; i < ell_a; i++)
      Ca[i + (size_t) j * lda] = T[i + (size_t) order[j] * ell];
  double *dC = (double *) dmalloc (sizeof (double) * (size_t) lda * ldk);
  if (!dC) return 5;
  rc = h2d (dC, Ca.data (), sizeof (double) * Ca.size ());
  if (!rc) rc = k_gemm_nn (Qa, 0, Pa, dC, lda, 1.0, 0.0, g->U, 0, Pk);
  if (!rc) rc = dev_sync ();
  dfree (dC);
  LGH_TRY (rc);
  if (Qb && ell > ell_a) {
    const int nb = ell - ell_a, Pb = npan (nb), ldb = Pb * 64;
    Cb.assign ((size_t) ldb * ldk, 0.0);
    for (int j = 0; j < k; j++)
      for (int i = 0; i < nb; i++)
        Cb[i + (size_t) j * ldb] = T[ell_a + i + (size_t) order[j] * ell];
    dC = (double *) dmalloc (sizeof (double) * (size_t) ldb * ldk);
    if (!dC) return 5;
    rc = h2d (dC, Cb.data (), sizeof (double) * Cb.size ());
    if (!rc) rc = k_gemm_nn (*Qb, 0, Pb, dC, ldb, 1.0, 1.0, g->U, 0, Pk);
    if (!rc) rc = dev_sync ();
    dfree (dC);
    LGH_TRY (rc);
  }
  return 0;
}

// residual gate (glr_common_impl.h:576-600): max_j ||F u_j - lam_raw_j u_j||
static int
residual_gate (lgh_glr_t *g, const std::vector<double> &evals,
               const std::vector<int> &order, int prec,
               lgh_glr_report_t *rep)
{
  const int k = g->kept;
  if (!g->opts.check || k == 0) return 0;
  PMat U, FU, Wk;
  const int Pk = npan (k);
  double *s = (double *) dmalloc (sizeof (double) * Pk * 64 * 2);
  if (!s) return 5;
  int rc = pmat_alloc (FU, g->n, k, 64, prec);
  if (!rc) rc = pmat_alloc (Wk, g->n, k, 64, prec);
  if (!rc && prec != F64) {
    rc = pmat_alloc (U, g->n, k, 64, prec);
    if (!rc) rc = k_copy_cols (g->U, 0, U, 0, k);
  }
  const PMat &Uin = (prec == F64) ? g->U : U;
  const double t0 = MPI_Wtime ();
  if (!rc) rc = apply_F (g, Uin, FU, Wk, pc_for (g, prec));
  std::vector<double> lr ((size_t) Pk * 64, 0.0), r2 ((size_t) Pk * 64, 0.0);
  for (int j = 0; j < k; j++) lr[j] = evals[order[j]];
  if (!rc) rc = h2d (s, lr.data (), sizeof (double) * Pk * 64);
  if (!rc) rc = k_colresid2 (FU, Uin, s, s + Pk * 64);
  if (!rc) rc = d2h (r2.data (), s + Pk * 64, sizeof (double) * Pk * 64);
  rep->t_operator += MPI_Wtime () - t0;
  for (int j = 0; j < k && !rc; j++)
    rep->resid_max = std::max (rep->resid_max, std::sqrt (r2[j]));
  pmat_free (U); pmat_free (FU); pmat_free (Wk);
  dfree (s);
  return rc;
}

static int
timed_F (lgh_glr_t *g, const PMat &X, PMat &Y, PMat &Wk, double *t_op,
         int pcp = -1)
{
  const double t0 = MPI_Wtime ();
  LGH_TRY (apply_F (g, X, Y, Wk, pcp));
  LGH_TRY (dev_sync ());
  *t_op += MPI_Wtime () - t0;
  return 0;
}

// panels q0 .. q0 + np - 1 of X as a matrix of its own (a view: never freed)
static PMat
pview (const PMat &X, int q0, int np)
{
  PMat v = X;
  v.p = X.panel (q0);
  v.ncols = std::min (X.ncols - X.W * q0, X.W * np);
  return v;
}

// panels solved together in one F sweep chunk: xopts.sweep_panels, or all
// P unless their work (the Z-solve buffers, about 64 B per element, and
// the sweep's scratch) would not fit in 85 % of device memory beside the
// blocks a build keeps (Q, Y and, for CholQR2, its copy: 24 B per element;
// 16 B with BCGS2)
static int
sweep_chunk (const lgh_glr_t *g, int P)
{
  if (g->x.sweep_panels > 0) return std::min (g->x.sweep_panels, P);
  const double total = (double) dmem_total ();
  if (total <= 0.0) return P;
  const double el = (double) g->n * 64.0;        // elements per panel
  const double keep = (g->implicit ? 16.0 : 24.0) * el * npan (g->ell > 0 ? g->ell : g->opts.ell);
  const double avail = 0.85 * total - keep, per_panel = 64.0 * el;
  if (avail >= per_panel * P) return P;
  return std::max (1, (int) (avail / per_panel));
}

// Y = F X, the panels of X in chunks (sweep_chunk); panels are independent,
// so the result does not depend on the chunking
static int
sweep_F (lgh_glr_t *g, const PMat &X, PMat &Y, double *t_op, int pcp = -1)
{
  const int P = X.npanels (), cp = sweep_chunk (g, P);
  PMat Wk;
  LGH_TRY (pmat_alloc (Wk, X.nrows, std::min (X.ncols, 64 * cp), 64, X.prec));
  int rc = 0;
  for (int q0 = 0; q0 < P && !rc; q0 += cp) {
    PMat Yc = pview (Y, q0, cp), Wc = Wk;
    const PMat Xc = pview (X, q0, cp);
    Wc.ncols = Xc.ncols;
    rc = timed_F (g, Xc, Yc, Wc, t_op, pcp);
  }
  pmat_free (Wk);
  return rc;
}

// sketch sweep Y64 = F Omega(seed, columns col0..col0+nc-1) at precision
// ps (FP32: vectors, B and Z in FP32; the V-cycle in FP32, or BF16 when
// that is the preconditioner precision), result in FP64; Omega is drawn
// chunk by chunk (sweep_chunk), so only one chunk of it is held
static int
sketch_sweep (lgh_glr_t *g, int ps, int64_t col0, int nc, PMat &Y64,
              double *t_op)
{
  const int n = g->n, P = npan (nc), cp = sweep_chunk (g, P);
  const int cc = std::min (nc, 64 * cp);
  PMat Om, Wk, Yp;
  LGH_TRY (pmat_alloc (Om, n, cc, 64, ps));
  int rc = pmat_alloc (Wk, n, cc, 64, ps);
  if (!rc && ps != F64) rc = pmat_alloc (Yp, n, cc, 64, ps);
  for (int q0 = 0; q0 < P && !rc; q0 += cp) {
    PMat O = Om, W = Wk, Yc = pview (Y64, q0, cp);
    O.ncols = W.ncols = Yc.ncols;
    // (k_randn writes the padding columns of a partial last panel as 0)
    rc = k_randn (O, g->opts.seed, 0, col0 + 64 * (int64_t) q0);
    if (rc) break;
    if (ps == F64) rc = timed_F (g, O, Yc, W, t_op, pc_for (g, F64));
    else {
      PMat Y = Yp;
      Y.ncols = O.ncols;
      rc = timed_F (g, O, Y, W, t_op, pc_for (g, ps));
      if (!rc) rc = k_copy_cols (Y, 0, Yc, 0, Y.ncols);
    }
  }
  pmat_free (Om); pmat_free (Wk); pmat_free (Yp);
  return rc;
}

// the build (glr_common_impl.h:415-608) under the precision policy of
// lgh_glr_xopts_t: every block after the sketch is FP64
static int
build (lgh_glr_t *g, lgh_glr_report_t *rep)
{
  const int n = g->n, ell = g->opts.ell, P = npan (ell);
  const int ps = to_prec (g->x.prec_sketch);
  const int pp = to_prec (g->x.prec_precond);
  const int pc = to_prec (g->x.prec_check);
  memset (rep, 0, sizeof (*rep));
  PMat Q, Y;
  LGH_TRY (pmat_alloc (Q, n, ell, 64, F64));
  LGH_TRY (pmat_alloc (Y, n, ell, 64, F64));
  // orthonormalization: CholQR2 of the whole block (replicated backend,
  // glr_common_impl.h:368-407) or BCGS2 in 64-column panels (the
  // implicit-U backend, glr_scalapack_impl.h:153-195 called at 576 and
  // 584, in place)
  auto orth = [&] () {
    return g->implicit ? bcgs2 (nullptr, 0, Y, ell) : cholqr (Y, ell, 2, nullptr);
  };
  int rc = sketch_sweep (g, ps, 0, ell, Y, &rep->t_operator);   // F Omega
  for (int p = 0; p < g->opts.q_power && !rc; p++) {
    double t0 = MPI_Wtime ();
    rc = orth ();
    if (!rc) rc = dev_sync ();
    rep->t_dense += MPI_Wtime () - t0;
    if (!rc) std::swap (Q.p, Y.p);                              // Q <- Y
    if (!rc) rc = sweep_F (g, Q, Y, &rep->t_operator, pp);
  }
  if (!rc) {
    double t0 = MPI_Wtime ();
    rc = orth ();
    if (!rc) std::swap (Q.p, Y.p);
    if (!rc) rc = dev_sync ();
    rep->t_dense += MPI_Wtime () - t0;
  }
  if (!rc) rc = sweep_F (g, Q, Y, &rep->t_operator, pp);      // Y = F Q
  std::vector<double> evals, Araw ((size_t) ell * ell);
  std::vector<int> order;
  if (!rc) {
    // T = Q^T Y (raw, kept for extension) — glr_common_impl.h:469-489
    const int ld = P * 64;
    double *dT = (double *) dmalloc (sizeof (double) * (size_t) ld * ld);
    std::vector<double> hT ((size_t) ld * ld);
    rc = dT ? 0 : 5;
    if (!rc) rc = k_gram (Q, 0, P, Y, 0, P, dT, ld, 0);
    if (!rc) rc = d2h (hT.data (), dT, sizeof (double) * hT.size ());
    dfree (dT);
    for (int j = 0; j < ell; j++)
      for (int i = 0; i < ell; i++) Araw[i + (size_t) j * ell] = hT[i + (size_t) j * ld];
  }
  pmat_free (Y);
  if (!rc) rc = rayleigh_ritz (g, ell, Araw, Q, ell, nullptr, rep, evals, order,
                               &rep->t_dense, g->implicit && g->opts.check);
  if (!rc && !g->implicit) rc = residual_gate (g, evals, order, pc, rep);
  if (!rc) {
    pmat_free (g->Q);
    g->Q = Q;
    Q.p = nullptr;
    g->Araw = Araw;
    g->ell = ell;
  }
  pmat_free (Q);
  return rc;
}

extern "C" int
lgh_glr_compute_x (Mat B, lgh_prior_t *prior, const lgh_glr_opts_t *opts,
                   const lgh_glr_xopts_t *xopts, lgh_glr_t **glr,
                   lgh_glr_report_t *report)
{
  PetscCheck (B != NULL && prior != NULL && glr != NULL, 0, PETSC_ERR_ARG_NULL,
              "lgh_glr_compute: NULL argument");
  lgh_glr_opts_t o = opts ? *opts : lgh_glr_opts_default ();
  const int N = B->m;
  // glr_common_impl.h:651-664
  PetscCheck (o.ell > 0 && o.ell <= N, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_compute: need 0 < ell (%d) <= N (%d)", o.ell, N);
  // DEFAULT resolves to REPLICATED, as in a reference build without
  // ScaLAPACK; SCALAPACK selects the implicit-U representation at
  // communicator size 1 (glr_common_impl.h:673-677)
  if (o.backend == LGH_GLR_DEFAULT) o.backend = LGH_GLR_REPLICATED;
  PetscCheck (B->m == B->n && B->m == prior->n, 0, PETSC_ERR_ARG_SIZ,
              "lgh_glr_compute: B (%d x %d) and prior (%d) disagree", B->m,
              B->n, prior->n);
  PetscCheck (B->assembled, 0, PETSC_ERR_ARG_WRONGSTATE,
              "lgh_glr_compute: B not assembled");
  const lgh_glr_xopts_t xo = xopts ? *xopts : lgh_glr_xopts_default ();
  auto vec_ok = [] (int q) { return q == LGH_PREC_F64 || q == LGH_PREC_F32; };
  auto pc_ok = [&] (int q) { return vec_ok (q) || q == LGH_PREC_BF16; };
  PetscCheck (vec_ok (xo.prec_sketch) && vec_ok (xo.prec_check)
              && pc_ok (xo.prec_precond) && pc_ok (xo.prec_sketch_precond)
              && xo.solve_blocked_min >= 0 && xo.sweep_panels >= 0,
              0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_compute_x: precision codes out of range");
  if (dev_init ()) return API_FAIL ("lgh_glr_compute");
  lgh_glr_t *g = new lgh_glr_t ();
  g->opts = o;
  g->x = xo;
  g->n = N;
  g->implicit = (o.backend == LGH_GLR_SCALAPACK);
  g->B = B;
  PetscObjectReference ((PetscObject) B);
  g->prior = prior;
  prior_ref (prior);
  lgh_glr_report_t rep;
  if (build (g, &rep)) {
    PetscErrorCode e = API_FAIL ("lgh_glr_compute");
    lgh_glr_destroy (g);
    return e;
  }
  if (report) *report = rep;
  *glr = g;
  return 0;
}

extern "C" int
lgh_glr_compute (Mat B, lgh_prior_t *prior, const lgh_glr_opts_t *opts,
                 lgh_glr_t **glr, lgh_glr_report_t *report)
{
  return lgh_glr_compute_x (B, prior, opts, nullptr, glr, report);
}

// incremental extension (glr_scalapack_impl.h:689-915), replicated T
static int
extend_incr (lgh_glr_t *g, int k_new, lgh_glr_report_t *rep)
{
  const int n = g->n, ell1 = g->ell, ell2 = ell1 + k_new;
  const int ps = to_prec (g->x.prec_sketch);
  const int pp = to_prec (g->x.prec_precond);
  const int pc = to_prec (g->x.prec_check);
  memset (rep, 0, sizeof (*rep));
  PMat Qn, Yn;
  LGH_TRY (pmat_alloc (Qn, n, k_new, 64, F64));
  LGH_TRY (pmat_alloc (Yn, n, k_new, 64, F64));
  // hashed columns continue the sequence past ell (:726-729)
  int rc = sketch_sweep (g, ps, ell1, k_new, Yn, &rep->t_operator);
  for (int p = 0; p < g->opts.q_power && !rc; p++) {          // :739-747
    double t0 = MPI_Wtime ();
    rc = bcgs2 (nullptr, 0, Yn, k_new);
    if (!rc) rc = dev_sync ();
    rep->t_dense += MPI_Wtime () - t0;
    if (!rc) std::swap (Qn.p, Yn.p);
    if (!rc) rc = sweep_F (g, Qn, Yn, &rep->t_operator, pp);
  }
  if (!rc) {                                                  // :749-768
    double t0 = MPI_Wtime ();
    rc = bcgs2 (&g->Q, ell1, Yn, k_new);
    if (!rc) rc = dev_sync ();
    rep->t_dense += MPI_Wtime () - t0;
    if (!rc) std::swap (Qn.p, Yn.p);
  }
  if (!rc) rc = sweep_F (g, Qn, Yn, &rep->t_operator, pp);      // :789
  // border Cnew = Q_all^T (F Q_new) (:792-814)
  std::vector<double> A2 ((size_t) ell2 * ell2, 0.0);
  if (!rc) {
    const int P1 = npan (ell1), Pn = npan (k_new);
    const int ld1 = P1 * 64, ldn = Pn * 64;
    double *top = (double *) dmalloc (sizeof (double) * (size_t) ld1 * ldn);
    double *bot = (double *) dmalloc (sizeof (double) * (size_t) ldn * ldn);
    std::vector<double> ht ((size_t) ld1 * ldn), hb ((size_t) ldn * ldn);
    rc = (top && bot) ? 0 : 5;
    if (!rc) rc = k_gram (g->Q, 0, P1, Yn, 0, Pn, top, ld1, 0);
    if (!rc) rc = k_gram (Qn, 0, Pn, Yn, 0, Pn, bot, ldn, 0);
    if (!rc) rc = d2h (ht.data (), top, sizeof (double) * ht.size ());
    if (!rc) rc = d2h (hb.data (), bot, sizeof (double) * hb.size ());
    dfree (top); dfree (bot);
    auto Cn = [&] (int i, int jn) {
      return i < ell1 ? ht[i + (size_t) jn * ld1] : hb[(i - ell1) + (size_t) jn * ldn];
    };
    // bordered raw A (:819-840): new rows of old columns by symmetry of F
    for (int j = 0; j < ell1; j++) {
      for (int i = 0; i < ell1; i++)
        A2[i + (size_t) j * ell2] = g->Araw[i + (size_t) j * ell1];
      for (int i = ell1; i < ell2; i++) A2[i + (size_t) j * ell2] = Cn (j, i - ell1);
    }
    for (int j = ell1; j < ell2; j++)
      for (int i = 0; i < ell2; i++) A2[i + (size_t) j * ell2] = Cn (i, j - ell1);
  }
  pmat_free (Yn);
  std::vector<double> evals;
  std::vector<int> order;
  const int old_ell = g->opts.ell;
  if (!rc) {
    g->opts.ell = ell2;
    rc = rayleigh_ritz (g, ell2, A2, g->Q, ell1, &Qn, rep, evals, order,
                        &rep->t_dense);
  }
  // (the implicit-U backend's extension runs no check:
  // glr_scalapack_impl.h:893-915)
  if (!rc && !g->implicit) rc = residual_gate (g, evals, order, pc, rep);
  if (!rc) {
    PMat Q2;
    rc = pmat_alloc (Q2, n, ell2, 64, F64);
    if (!rc) rc = k_copy_cols (g->Q, 0, Q2, 0, ell1);
    if (!rc) rc = k_copy_cols (Qn, 0, Q2, ell1, k_new);
    if (!rc) rc = dev_sync ();
    if (!rc) {
      pmat_free (g->Q);
      g->Q = Q2;
      g->Araw = A2;
      g->ell = ell2;
    }
    else pmat_free (Q2);
  }
  if (rc) g->opts.ell = old_ell;
  pmat_free (Qn);
  return rc;
}

extern "C" int
lgh_glr_extend (lgh_glr_t *glr, int k_new, lgh_glr_report_t *report)
{
  // glr_common_impl.h:695-704
  PetscCheck (!glr->corrected, 0, PETSC_ERR_ORDER,
              "lgh_glr_extend: not available after lgh_glr_correct");
  PetscCheck (k_new > 0, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_extend: k_new must be positive");
  PetscCheck (glr->opts.ell + k_new <= glr->n, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_extend: extended ell (%d) exceeds N (%d)",
              glr->opts.ell + k_new, glr->n);
  lgh_glr_report_t rep;
  if (extend_incr (glr, k_new, &rep)) return API_FAIL ("lgh_glr_extend");
  if (report) *report = rep;
  return 0;
}

extern "C" void
lgh_glr_destroy (lgh_glr_t *g)
{
  if (g == NULL) return;
  pmat_free (g->U);
  pmat_free (g->Q);
  dfree (g->dV);
  MatDestroy (&g->B);
  lgh_prior_destroy (g->prior);
  delete g;
}

// ---- downstream -------------------------------------------------------------

static int
width_for (int m)
{
  int W = 1;
  while (W < m && W < 64) W *= 2;
  return W;
}

// Y = f_c X + U diag(f_lam - f_c) U^T X per column j (block master formula,
// glr_common_impl.h:909-966).  X, Y: W = 64 fp64 panels of m columns;
// fk[k + j*kept] = f_lam_k(j) - f_c(j), fc[j] = f_c(j).
static int
woodbury (lgh_glr_t *g, const PMat &X, int m, const std::vector<double> &fk,
          const std::vector<double> &fc, PMat &Y)
{
  const int Pm = npan (m), ldm = Pm * 64;
  double *dfc = (double *) dmalloc (sizeof (double) * ldm);
  if (!dfc) return 5;
  std::vector<double> hfc ((size_t) ldm, 0.0);
  for (int j = 0; j < m; j++) hfc[j] = fc[j];
  int rc = h2d (dfc, hfc.data (), sizeof (double) * ldm);
  if (!rc) rc = k_colscale (dfc, X, Y);
  dfree (dfc);
  LGH_TRY (rc);
  const int k = g->kept;
  if (k == 0) return 0;
  const int Pk = npan (k), ldk = Pk * 64;
  if (g->implicit) {
    // U^T X = V^T (Q^T X) and U W = Q (V W) (glr_scalapack_impl.h:923-1027):
    // the two tall products over Q on the MFMA kernels, the small ones in
    // rocBLAS, the filter values applied on the host
    const int Pl = npan (g->ell), ldl = Pl * 64;
    double *G = (double *) dmalloc (sizeof (double) * (size_t) ldl * ldm);
    double *W = (double *) dmalloc (sizeof (double) * (size_t) ldk * ldm);
    rc = (G && W) ? 0 : 5;
    std::vector<double> hw ((size_t) ldk * ldm);
    if (!rc) rc = k_gram (g->Q, 0, Pl, X, 0, Pm, G, ldl, 0);
    if (!rc) rc = blas_dgemm (1, 0, ldk, ldm, ldl, 1.0, g->dV, ldl, G, ldl, 0.0, W, ldk);
    if (!rc) rc = d2h (hw.data (), W, sizeof (double) * hw.size ());
    if (!rc) {
      for (int j = 0; j < ldm; j++)
        for (int i = 0; i < ldk; i++) {
          double &w = hw[i + (size_t) j * ldk];
          w = (i < k && j < m) ? w * fk[i + (size_t) j * k] : 0.0;
        }
      rc = h2d (W, hw.data (), sizeof (double) * hw.size ());
    }
    if (!rc) rc = blas_dgemm (0, 0, ldl, ldm, ldk, 1.0, g->dV, ldl, W, ldk, 0.0, G, ldl);
    if (!rc) rc = k_gemm_nn (g->Q, 0, Pl, G, ldl, 1.0, 1.0, Y, 0, Pm);
    if (!rc) rc = dev_sync ();
    dfree (G); dfree (W);
    return rc;
  }
  double *W = (double *) dmalloc (sizeof (double) * (size_t) ldk * ldm);
  if (!W) return 5;
  std::vector<double> hw ((size_t) ldk * ldm);
  rc = k_gram (g->U, 0, Pk, X, 0, Pm, W, ldk, 0);
  if (!rc) rc = d2h (hw.data (), W, sizeof (double) * hw.size ());
  if (!rc) {
    for (int j = 0; j < ldm; j++)
      for (int i = 0; i < ldk; i++) {
        double &w = hw[i + (size_t) j * ldk];
        w = (i < k && j < m) ? w * fk[i + (size_t) j * k] : 0.0;
      }
    rc = h2d (W, hw.data (), sizeof (double) * hw.size ());
  }
  if (!rc) rc = k_gemm_nn (g->U, 0, Pk, W, ldk, 1.0, 1.0, Y, 0, Pm);
  if (!rc) rc = dev_sync ();
  dfree (W);
  return rc;
}

// block driver for the full-space operations: op selects
//   0 solve   : x = Z^{-T} M^{1/2} (F+cI)^{-1} M^{1/2} Z^{-1} rhs
//   3 GINVT   : x = Z^{-T} M^{1/2} (F+cI)^{-1/2} in   (== sample)
//   2 GINV    : x = (F+cI)^{-1/2} M^{1/2} Z^{-1} in
//   0x10 G    : x = Z M^{-1/2} (F+cI)^{1/2} in
//   0x11 GT   : x = (F+cI)^{1/2} M^{-1/2} Z^T in
//   0x20 pow  : x = (F+cI)^p in (prior-weighted frame)
//   0x30 filter with caller values
struct OpSpec {
  int kind;
  double p;
  const double *f_lam;
  double f_c;
};

static void
coeffs (lgh_glr_t *g, const OpSpec &op, int m, const double *c,
        std::vector<double> &fk, std::vector<double> &fc)
{
  const int k = g->kept;
  fk.assign ((size_t) std::max (k, 1) * m, 0.0);
  fc.assign (m, 0.0);
  const double half = 0.5;
  for (int j = 0; j < m; j++) {
    const double cj = c ? c[j] : 0.0;
    double f_c;
    for (int i = 0; i < k; i++) {
      const double l = g->lam[i];
      double f;
      switch (op.kind) {
      case 0: f = 1. / (l + cj); break;                         // :1018-1019
      case 2: case 3: f = 1. / std::pow (l + cj, half); break;  // :1079-1086
      case 0x10: case 0x11: f = std::pow (l + cj, half); break; // :1063-1073
      case 0x20: f = std::pow (l + cj, op.p); break;            // :983-984
      default: f = op.f_lam[i]; break;
      }
      fk[i + (size_t) j * k] = f;
    }
    switch (op.kind) {
    case 0: f_c = 1. / cj; break;
    case 2: case 3: f_c = 1. / std::sqrt (cj); break;
    case 0x10: case 0x11: f_c = std::sqrt (cj); break;
    case 0x20: f_c = std::pow (cj, op.p); break;
    default: f_c = op.f_c; break;
    }
    fc[j] = f_c;
    for (int i = 0; i < k; i++) fk[i + (size_t) j * k] -= f_c;
  }
}

// tier: -1 by the policy (solve_blocked_min), 0 per-column PCG (the
// reference's vector path), 1 the blocked tier (assembled Z, modes >= 2).
// An explicit tier also fixes the panel width at 64, so that a column's
// arithmetic does not depend on how many columns share the call.
static int
run_op (lgh_glr_t *g, const OpSpec &op, int m, const double *c,
        const double *in, int ldi, double *out, int ldo, int tier = -1,
        bool dev = false)
{
  lgh_prior_t *p = g->prior;
  const bool can_block = p->is_mat && p->o.blocked_mode >= 2;
  const bool fixed_w = tier >= 0;
  if (tier < 0)
    tier = (g->x.solve_blocked_min > 0 && m >= g->x.solve_blocked_min) ? 1 : 0;
  const bool blocked = tier == 1 && can_block;
  const int n = g->n, Wz = (blocked || fixed_w) ? 64 : width_for (m);
  auto zsolve = [&] (int trans, const PMat &X, PMat &Y) -> int {
    if (blocked) return prior_solve_block (p, trans, X, Y, pc_for (g, F64));
    return prior_solve_vec (p, trans, X, Y);
  };
  LGH_TRY (prior_setup_device (p, F64));
  std::vector<double> fk, fc;
  coeffs (g, op, m, c, fk, fc);
  PMat H, A, Bm, X64, Y64;
  // contiguous host blocks (ld == n) move directly; others through hb;
  // device blocks (dev) by pitched device copies
  std::vector<double> hb;
  const double *hin = in;
  if (!dev && (ldi != n || ldo != n)) hb.resize ((size_t) n * m);
  if (!dev && ldi != n) {
    for (int j = 0; j < m; j++) memcpy (&hb[(size_t) j * n], in + (size_t) j * ldi, sizeof (double) * n);
    hin = hb.data ();
  }
  const size_t colb = sizeof (double) * (size_t) n;
  int rc = pmat_alloc (H, n, m, 1, F64);
  if (!rc) rc = pmat_alloc (A, n, m, Wz, F64);
  if (!rc) rc = pmat_alloc (Bm, n, m, Wz, F64);
  if (!rc) rc = pmat_alloc (X64, n, m, 64, F64);
  if (!rc) rc = pmat_alloc (Y64, n, m, 64, F64);
  if (!rc) rc = dev ? dcopy2d (H.p, colb, in, sizeof (double) * (size_t) ldi, colb, m)
                    : h2d (H.p, hin, colb * m);
  if (!rc) rc = k_copy_cols (H, 0, A, 0, m);
  // left part (before the filter): into X64
  if (!rc) {
    switch (op.kind) {
    case 0: case 2:                    // M^{1/2} Z^{-1} in
      rc = zsolve (0, A, Bm);
      if (!rc) rc = k_diag_scale (Bm, p->d_msqrt, Bm);
      if (!rc) rc = k_copy_cols (Bm, 0, X64, 0, m);
      break;
    case 0x11:                         // M^{-1/2} Z^T in
      rc = prior_apply_Z (p, 1, A, Bm);
      if (!rc) rc = k_diag_scale (Bm, p->d_minvsqrt, Bm);
      if (!rc) rc = k_copy_cols (Bm, 0, X64, 0, m);
      break;
    default:
      rc = k_copy_cols (A, 0, X64, 0, m);
      break;
    }
  }
  if (!rc) rc = woodbury (g, X64, m, fk, fc, Y64);
  if (!rc) rc = k_copy_cols (Y64, 0, A, 0, m);
  if (!rc) {
    switch (op.kind) {
    case 0: case 3:                    // Z^{-T} M^{1/2} (.)
      rc = k_diag_scale (A, p->d_msqrt, A);
      if (!rc) rc = zsolve (1, A, Bm);
      if (!rc) rc = k_copy_cols (Bm, 0, H, 0, m);
      break;
    case 0x10:                         // Z M^{-1/2} (.)
      rc = k_diag_scale (A, p->d_minvsqrt, A);
      if (!rc) rc = prior_apply_Z (p, 0, A, Bm);
      if (!rc) rc = k_copy_cols (Bm, 0, H, 0, m);
      break;
    default:
      rc = k_copy_cols (A, 0, H, 0, m);
      break;
    }
  }
  double *hout = (ldo == n || dev) ? out : hb.data ();
  if (!rc && dev) {
    rc = dcopy2d (out, sizeof (double) * (size_t) ldo, H.p, colb, colb, m);
    if (!rc) rc = dev_sync ();
  } else if (!rc)
    rc = d2h (hout, H.p, colb * m);
  pmat_free (H); pmat_free (A); pmat_free (Bm); pmat_free (X64); pmat_free (Y64);
  LGH_TRY (rc);
  if (!dev && ldo != n)
    for (int j = 0; j < m; j++) memcpy (out + (size_t) j * ldo, &hb[(size_t) j * n], sizeof (double) * n);
  return 0;
}

int
lgh::glr_filter_panels (lgh_glr_t *g, const double *f_lam, double f_c,
                        const PMat &X, int m, PMat &Y)
{
  const int k = g->kept;
  std::vector<double> fk ((size_t) std::max (k, 1) * m, 0.0), fc (m, f_c);
  for (int j = 0; j < m; j++)
    for (int i = 0; i < k; i++) fk[i + (size_t) j * k] = f_lam[i] - f_c;
  return woodbury (g, X, m, fk, fc, Y);
}

int
lgh::glr_run_block (lgh_glr_t *g, int op, int m, const double *c,
                    const double *in, int ldi, double *out, int ldo, int tier)
{
  LGH_CHECK (op == 0 || op == 3, 8, "glr_run_block: op must be solve or sample");
  return run_op (g, OpSpec { op, 0, nullptr, 0 }, m, c, in, ldi, out, ldo, tier);
}

static PetscErrorCode
vec_op (lgh_glr_t *g, const OpSpec &op, double c, Vec in, Vec out,
        const char *name)
{
  PetscCheck (in->n == g->n && out->n == g->n, 0, PETSC_ERR_ARG_SIZ,
              "%s: vector size mismatch", name);
  if (run_op (g, op, 1, &c, in->a, g->n, out->a, g->n)) return API_FAIL (name);
  return 0;
}

extern "C" int
lgh_glr_solve (lgh_glr_t *g, double c, Vec rhs, Vec x)
{
  PetscCheck (c > g->floor, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_solve: shift c must exceed the spectrum floor %g", g->floor);
  return vec_op (g, OpSpec { 0, 0, nullptr, 0 }, c, rhs, x, "lgh_glr_solve");
}

extern "C" int
lgh_glr_factor (lgh_glr_t *g, double c, lgh_factor_t which, Vec in, Vec out)
{
  PetscCheck (c > g->floor, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_factor: shift c must exceed the spectrum floor %g", g->floor);
  int kind;
  switch (which) {
  case LGH_FACTOR_G: kind = 0x10; break;
  case LGH_FACTOR_GT: kind = 0x11; break;
  case LGH_FACTOR_GINV: kind = 2; break;
  case LGH_FACTOR_GINVT: kind = 3; break;
  default:
    SETERRQ (0, PETSC_ERR_ARG_OUTOFRANGE, "lgh_glr_factor: unknown factor action");
  }
  return vec_op (g, OpSpec { kind, 0, nullptr, 0 }, c, in, out, "lgh_glr_factor");
}

extern "C" int
lgh_glr_sample (lgh_glr_t *g, double c, Vec xi, Vec x)
{
  return lgh_glr_factor (g, c, LGH_FACTOR_GINVT, xi, x);     // :1099-1104
}

extern "C" int
lgh_glr_pow (lgh_glr_t *g, double c, double p, Vec in, Vec out)
{
  PetscCheck (c > g->floor, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_glr_pow: shift c must exceed the spectrum floor %g", g->floor);
  return vec_op (g, OpSpec { 0x20, p, nullptr, 0 }, c, in, out, "lgh_glr_pow");
}

extern "C" int
lgh_glr_filter (lgh_glr_t *g, const double *f_lam, double f_c, Vec in, Vec out)
{
  return vec_op (g, OpSpec { 0x30, 0, f_lam, f_c }, 0.0, in, out,
                 "lgh_glr_filter");
}

extern "C" double
lgh_glr_logdet (lgh_glr_t *g, double c)
{
  double s = 0.;                                              // :988-997
  for (int i = 0; i < g->kept; i++) s += std::log (g->lam[i] + c) - std::log (c);
  return s;
}

extern "C" int
lgh_glr_eigs (const lgh_glr_t *g, int *nkept, const double **lam)
{
  *nkept = g->kept;
  *lam = g->lam.data ();
  return 0;
}

extern "C" int
lgh_glr_apply (lgh_glr_t *g, double c, Vec v, Vec y)
{
  // exact y = B v + c Z M^{-1} Z^T v (glr_common_impl.h:1043-1047)
  PetscCheck (v->n == g->n && y->n == g->n, 0, PETSC_ERR_ARG_SIZ,
              "lgh_glr_apply: vector size mismatch");
  lgh_prior_t *p = g->prior;
  if (g->corrected) {
    // after a correction B alone no longer represents the operator: the
    // corrected low-rank action Z M^{-1/2} (U lam U^T + cI) M^{-1/2} Z^T v
    // (glr_common_impl.h:1029-1041)
    PMat X, T;
    std::vector<double> f (std::max (g->kept, 1));
    for (int i = 0; i < g->kept; i++) f[i] = g->lam[i] + c;
    int rc = prior_setup_device (p, F64);
    if (!rc) rc = pmat_alloc (X, g->n, 1, 64, F64);
    if (!rc) rc = pmat_alloc (T, g->n, 1, 64, F64);
    PMat H;
    if (!rc) rc = pmat_alloc (H, g->n, 1, 1, F64);
    if (!rc) rc = h2d (H.p, v->a, sizeof (double) * g->n);
    if (!rc) rc = k_copy_cols (H, 0, X, 0, 1);
    if (!rc) rc = prior_apply_Z (p, 1, X, T);
    if (!rc) rc = k_diag_scale (T, p->d_minvsqrt, T);
    if (!rc) rc = glr_filter_panels (g, f.data (), c, T, 1, X);
    if (!rc) rc = k_diag_scale (X, p->d_minvsqrt, X);
    if (!rc) rc = prior_apply_Z (p, 0, X, T);
    if (!rc) rc = k_copy_cols (T, 0, H, 0, 1);
    if (!rc) rc = d2h (y->a, H.p, sizeof (double) * g->n);
    pmat_free (X); pmat_free (T); pmat_free (H);
    if (rc) return API_FAIL ("lgh_glr_apply");
    return 0;
  }
  PMat X, T, Yb;
  DevCSR *dB = nullptr;
  int rc = prior_setup_device (p, F64);
  if (!rc) rc = mat_device_csr (g->B, F64, 0, &dB);
  if (!rc) rc = pmat_alloc (X, g->n, 1, 1, F64);
  if (!rc) rc = pmat_alloc (T, g->n, 1, 1, F64);
  if (!rc) rc = pmat_alloc (Yb, g->n, 1, 1, F64);
  if (!rc) rc = h2d (X.p, v->a, sizeof (double) * g->n);
  if (!rc) rc = k_spmm (*dB, 1.0, X, 0.0, Yb);
  if (!rc) rc = prior_apply_Z (p, 1, X, T);
  if (!rc) rc = k_diag_scale (T, p->d_minv, T);
  if (!rc) rc = prior_apply_Z (p, 0, T, X);
  if (!rc) rc = k_axpby (c, X, 1.0, Yb);
  if (!rc) rc = d2h (y->a, Yb.p, sizeof (double) * g->n);
  pmat_free (X); pmat_free (T); pmat_free (Yb);
  if (rc) return API_FAIL ("lgh_glr_apply");
  return 0;
}

extern "C" int
lgh_glr_solve_block (lgh_glr_t *g, int m, const double *c, const double *rhs,
                     int ldr, double *x, int ldx)
{
  for (int j = 0; j < m; j++)
    PetscCheck (c[j] > g->floor, 0, PETSC_ERR_ARG_OUTOFRANGE,
                "lgh_glr_solve_block: shift c[%d] must exceed %g", j, g->floor);
  if (run_op (g, OpSpec { 0, 0, nullptr, 0 }, m, c, rhs, ldr, x, ldx))
    return API_FAIL ("lgh_glr_solve_block");
  return 0;
}

extern "C" int
lgh_glr_sample_block (lgh_glr_t *g, int m, const double *c, const double *xi,
                      int ldr, double *x, int ldx)
{
  for (int j = 0; j < m; j++)
    PetscCheck (c[j] > g->floor, 0, PETSC_ERR_ARG_OUTOFRANGE,
                "lgh_glr_sample_block: shift c[%d] must exceed %g", j, g->floor);
  if (run_op (g, OpSpec { 3, 0, nullptr, 0 }, m, c, xi, ldr, x, ldx))
    return API_FAIL ("lgh_glr_sample_block");
  return 0;
}

// the two block operations on device-resident blocks: the same run_op
// (same kernels, same bits per column) with pitched device copies in place
// of the host transfers
static PetscErrorCode
block_dev (lgh_glr_t *g, int kind, int m, const double *c, const double *in,
           int ldi, double *out, int ldo, const char *name)
{
  PetscCheck (m >= 0 && ldi >= g->n && ldo >= g->n, 0, PETSC_ERR_ARG_SIZ,
              "%s: need m >= 0 and leading dimensions >= n = %d", name, g->n);
  if (m == 0) return 0;
  PetscCheck (c && in && out, 0, PETSC_ERR_ARG_NULL, "%s: null argument", name);
  PetscCheck (lgh::is_device_ptr (in) && lgh::is_device_ptr (out), 0,
              PETSC_ERR_ARG_WRONG, "%s: rhs and x must be device memory", name);
  // the blocks' last columns end inside their allocations
  const size_t span_in = sizeof (double) * ((size_t) ldi * (m - 1) + g->n);
  const size_t span_out = sizeof (double) * ((size_t) ldo * (m - 1) + g->n);
  PetscCheck (lgh::device_range_ok (in, span_in) && lgh::device_range_ok (out, span_out),
              0, PETSC_ERR_ARG_SIZ,
              "%s: an n x %d block with these leading dimensions extends past "
              "its device allocation", name, m);
  for (int j = 0; j < m; j++)
    PetscCheck (c[j] > g->floor, 0, PETSC_ERR_ARG_OUTOFRANGE,
                "%s: shift c[%d] must exceed %g", name, j, g->floor);
  if (run_op (g, OpSpec { kind, 0, nullptr, 0 }, m, c, in, ldi, out, ldo, -1, true))
    return API_FAIL (name);
  return 0;
}

extern "C" int
lgh_glr_solve_block_dev (lgh_glr_t *g, int m, const double *c,
                         const double *rhs, int ldr, double *x, int ldx)
{
  return block_dev (g, 0, m, c, rhs, ldr, x, ldx, "lgh_glr_solve_block_dev");
}

extern "C" int
lgh_glr_sample_block_dev (lgh_glr_t *g, int m, const double *c,
                          const double *xi, int ldr, double *x, int ldx)
{
  return block_dev (g, 3, m, c, xi, ldr, x, ldx, "lgh_glr_sample_block_dev");
}

// U = Q V of an implicit object into a new n x kept matrix
// (glr_scalapack_impl.h:1029-1076)
static int
form_U (const lgh_glr_t *g, PMat &U)
{
  const int k = g->kept, Pk = npan (std::max (k, 1)), Pl = npan (g->ell);
  LGH_TRY (pmat_alloc (U, g->n, std::max (k, 1), 64, F64));
  int rc = (k > 0) ? k_gemm_nn (g->Q, 0, Pl, g->dV, Pl * 64, 1.0, 0.0, U, 0, Pk) : 0;
  if (!rc) rc = dev_sync ();
  if (rc) pmat_free (U);
  return rc;
}

int
lgh::glr_materialize_U (lgh_glr_t *g)
{
  if (!g->implicit) return 0;
  PMat U;
  LGH_TRY (form_U (g, U));
  pmat_free (g->U);
  g->U = U;
  g->implicit = 0;
  g->V.clear ();
  dfree (g->dV);
  g->dV = nullptr;
  return 0;
}

extern "C" int
lgh_glr_get_U (const lgh_glr_t *g, double *U, int ldu)
{
  const int k = g->kept;
  if (k == 0) return 0;
  PMat H, Uf;
  std::vector<double> h ((size_t) g->n * k);
  int rc = g->implicit ? form_U (g, Uf) : 0;
  if (!rc) rc = pmat_alloc (H, g->n, k, 1, F64);
  if (!rc) rc = k_copy_cols (g->implicit ? Uf : g->U, 0, H, 0, k);
  if (!rc) rc = d2h (h.data (), H.p, sizeof (double) * h.size ());
  pmat_free (H);
  pmat_free (Uf);
  if (rc) return API_FAIL ("lgh_glr_get_U");
  for (int j = 0; j < k; j++) memcpy (U + (size_t) j * ldu, &h[(size_t) j * g->n], sizeof (double) * g->n);
  return 0;
}
