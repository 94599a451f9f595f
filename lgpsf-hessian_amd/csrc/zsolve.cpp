// zsolve.cpp — device Z-solve tiers (see zsolve.hpp).
#include "zsolve.hpp"

#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

namespace lgh {

DevMG::~DevMG ()
{
  free_work ();
  for (auto *a : A) delete a;
  for (auto *a : P) delete a;
  for (auto *a : PT) delete a;
  dfree (cinv);
  mgc_img.release ();
}

void
DevMG::Work::free ()
{
  for (auto *vv : { &R1, &R2, &D1, &D2, &Bl, &Xl })
    for (auto &m : *vv) pmat_free (m);
  R1.clear (); R2.clear (); D1.clear (); D2.clear (); Bl.clear (); Xl.clear ();
  np = -1;
}

void
DevMG::free_work ()
{
  for (Work &w : work) w.free ();
}

// Chebyshev recurrence scalars of zsolve_impl.h:113-138
struct ChebCoef {
  double theta, delta, sigma1, rho_old;
  ChebCoef (double lmin, double lmax)
  {
    theta = 0.5 * (lmax + lmin);
    delta = 0.5 * (lmax - lmin);
    sigma1 = theta / delta;
    rho_old = 1.0 / sigma1;
  }
  // (a, b) for step k (k = 1: a = 0, b = 1/theta)
  void step (int k, double *a, double *b)
  {
    if (k == 1) { *a = 0.0; *b = 1.0 / theta; return; }
    const double rho = 1.0 / (2.0 * sigma1 - rho_old);
    *a = rho * rho_old;
    *b = 2.0 * rho / delta;
    rho_old = rho;
  }
};

// the one-pass degree-2 smoother and explicit residuals serve the row
// kernels' widths
static inline bool
fused_level (const DevMG &g, int l, int W)
{
  return g.nu[l] == 2 && (W == 64 || W == 16 || W == 8);
}

int
mg_upload (const AmgHierarchy &h, int prec, DevMG **out)
{
  DevMG *g = new DevMG ();
  g->prec = prec;
  // BF16 is a vector storage precision: the matrices and the coarse inverse
  // of that tier are fp32
  const int mprec = (prec == BF16) ? F32 : prec;
  g->nl = (int) h.lev.size ();
  g->gamma = h.gamma;
  g->A.assign (g->nl, nullptr);
  g->P.assign (g->nl, nullptr);
  g->PT.assign (g->nl, nullptr);
  int rc = 0;
  for (int l = 0; l < g->nl && !rc; l++) {
    const AmgLevel &lv = h.lev[l];
    g->n.push_back (lv.A.n);
    g->slo.push_back (lv.slo);
    g->shi.push_back (lv.shi);
    g->nu.push_back (lv.nu);
    rc = csr_upload (lv.A.n, lv.A.m, lv.A.rp.data (), lv.A.ci.data (),
                     lv.A.v.data (), mprec, 1, &g->A[l]);
    if (!rc && lv.nu == 2 && l > 0) {
      // the one-pass smoother's first-step factor of this level
      double a1, b1;
      ChebCoef (lv.slo, lv.shi).step (1, &a1, &b1);
      rc = csr_scale_cols (*g->A[l], lv.A.rp.data (), lv.A.ci.data (),
                           lv.A.v.data (), b1);
    }
    if (!rc && l > 0) {
      rc = csr_upload (lv.P.n, lv.P.m, lv.P.rp.data (), lv.P.ci.data (),
                       lv.P.v.data (), mprec, 0, &g->P[l]);
      if (!rc) {
        HostCSR T = host_transpose (lv.P);
        rc = csr_upload (T.n, T.m, T.rp.data (), T.ci.data (), T.v.data (),
                         mprec, 0, &g->PT[l]);
      }
    }
  }
  if (!rc) {
    g->n0 = h.lev[0].A.n;
    const size_t cnt = (size_t) g->n0 * g->n0;
    g->cinv = dmalloc (cnt * prec_bytes (mprec));
    if (!g->cinv) rc = 5;
    else if (mprec == F64) rc = h2d (g->cinv, h.coarse_inv.data (), cnt * 8);
    else {
      std::vector<float> f (h.coarse_inv.begin (), h.coarse_inv.end ());
      rc = h2d (g->cinv, f.data (), cnt * 4);
    }
  }
  if (rc) { delete g; return rc; }
  *out = g;
  return 0;
}

static thread_local SolveWork *t_solve_work = nullptr;

SolveWork *solve_work () { return t_solve_work; }
SolveWorkScope::SolveWorkScope (SolveWork *w) : prev (t_solve_work) { t_solve_work = w; }
SolveWorkScope::~SolveWorkScope () { t_solve_work = prev; }

DevMG::Work *
SolveWork::mg_slots (const DevMG *g)
{
  for (auto &e : mg)
    if (e.first == g) return e.second;
  DevMG::Work *w = new DevMG::Work[3];
  mg.emplace_back (g, w);
  return w;
}

SolveWork::~SolveWork ()
{
  for (auto &e : mg) {
    for (int k = 0; k < 3; k++) e.second[k].free ();
    delete[] e.second;
  }
}

static int
mg_work (DevMG &g, const PMat &like, DevMG::Work **out)
{
  const int np = like.npanels ();
  const int slot = like.W == 64 ? 0 : like.W == 16 ? 1 : 2;
  DevMG::Work &w = t_solve_work ? t_solve_work->mg_slots (&g)[slot] : g.work[slot];
  *out = &w;
  if (w.W == like.W && w.np == np) return 0;
  w.free ();
  w.R1.resize (g.nl); w.R2.resize (g.nl); w.D1.resize (g.nl);
  w.D2.resize (g.nl); w.Bl.resize (g.nl); w.Xl.resize (g.nl);
  const int nc = np * like.W;
  for (int l = 0; l < g.nl; l++) {
    if (l > 0) {
      LGH_TRY (pmat_alloc (w.R1[l], g.n[l], nc, like.W, g.prec));
      if (!fused_level (g, l, like.W)) {        // the per-step smoother path
        LGH_TRY (pmat_alloc (w.R2[l], g.n[l], nc, like.W, g.prec));
        LGH_TRY (pmat_alloc (w.D1[l], g.n[l], nc, like.W, g.prec));
        LGH_TRY (pmat_alloc (w.D2[l], g.n[l], nc, like.W, g.prec));
      }
    }
    if (l < g.nl - 1) {
      LGH_TRY (pmat_alloc (w.Bl[l], g.n[l], nc, like.W, g.prec));
      LGH_TRY (pmat_alloc (w.Xl[l], g.n[l], nc, like.W, g.prec));
    }
  }
  w.W = like.W;
  w.np = np;
  return 0;
}

// correction-form smoothing of level l (zsolve_impl.h:385-404).
// x_zero: X is zero on entry (pre-smoothing), so R = B without a product.
// *Rout_final receives the buffer holding the recurrence residual B - A X
// after the last step.
static int
mg_smooth (DevMG &g, DevMG::Work &w, int l, const PMat &B, PMat &X,
           int x_zero, const PMat **Rfinal)
{
  const DevCSR &A = *g.A[l];
  ChebCoef cc (g.slo[l], g.shi[l]);
  PMat *Rin = &w.R1[l], *Rout = &w.R2[l];
  PMat *Dcur = nullptr, *Dn = &w.D1[l];
  const PMat *R0 = &B;
  if (!x_zero) {
    LGH_TRY (k_resid (A, B, X, w.R1[l]));
    R0 = &w.R1[l];
    Rout = &w.R2[l];
  }
  else Rout = &w.R1[l];
  for (int k = 1; k <= g.nu[l]; k++) {
    double a, b;
    cc.step (k, &a, &b);
    const PMat *Rsrc = (k == 1) ? R0 : Rin;
    const int xmode = (k == 1) ? (x_zero ? 0 : 1) : 1;
    LGH_TRY (k_cheb_step (A, 1, nullptr, a, b, Dcur, *Rsrc, X, xmode, *Dn, *Rout));
    // rotate: the output residual/direction become the inputs
    Rin = Rout;
    Rout = (Rin == &w.R1[l]) ? &w.R2[l] : &w.R1[l];
    Dcur = Dn;
    Dn = (Dcur == &w.D1[l]) ? &w.D2[l] : &w.D1[l];
  }
  if (Rfinal) *Rfinal = (g.nu[l] >= 1) ? Rin : R0;
  return 0;
}

// ---- the fused coarse cycle (k_mg_coarse) ----
// Levels 1 .. top run in one launch when each has at most g_mgc_rows rows
// and g_mgc_nnz nonzeros (0 rows: never); the finest level never does.
static std::atomic<int> g_mgc_rows { 1024 }, g_mgc_nnz { 65536 };

void
mg_coarse_config (int max_rows, int max_nnz)
{
  g_mgc_rows = max_rows < 0 ? 1024 : max_rows;
  g_mgc_nnz = max_nnz < 0 ? 65536 : max_nnz;
}

// the operations of mg_run (l, x_init) below, in order
static void
mgc_prog (const DevMG &g, int l, int x_init, std::vector<unsigned char> &ops)
{
  if (l == 0) { ops.push_back ((unsigned char) (MGC_DENSE << 4)); return; }
  auto op = [&] (int k) { ops.push_back ((unsigned char) (k << 4 | l)); };
  if (x_init) { op (MGC_RESID); op (MGC_SMX); }
  else op (MGC_SM0);
  op (MGC_RESID);
  op (MGC_RESTRICT);
  mgc_prog (g, l - 1, 0, ops);
  if (g.gamma == 2 && l - 1 >= 1) mgc_prog (g, l - 1, 1, ops);
  op (MGC_PROLONG);
  op (MGC_RESID);
  op (MGC_SMX);
}

// the coarse correction of level l as one fused launch when levels
// 0 .. l - 1 qualify (*done = 1), else nothing (*done = 0)
static int
mg_coarse_fused (DevMG &g, DevMG::Work &w, int l, int *done)
{
  *done = 0;
  const int top = l - 1;
  if (top < 1 || top >= MGC_MAX_LEV || g_mgc_rows.load () <= 0) return 0;
  for (int j = 1; j <= top; j++)
    if (!fused_level (g, j, w.W) || g.n[j] > g_mgc_rows.load ()
        || g.A[j]->nnz > g_mgc_nnz.load ())
      return 0;
  MgcLevelIn lev[MGC_MAX_LEV];
  for (int j = 0; j <= top; j++) {
    lev[j].B = &w.Bl[j];
    lev[j].X = &w.Xl[j];
    if (j == 0) continue;
    lev[j].R = &w.R1[j];
    lev[j].A = g.A[j];
    lev[j].P = g.P[j];
    lev[j].PT = g.PT[j];
    ChebCoef cc (g.slo[j], g.shi[j]);
    double a1;
    cc.step (1, &a1, &lev[j].b1);
    cc.step (2, &lev[j].a2, &lev[j].b2);
  }
  if (!mgc_supported (top + 1, lev, g.n0)) return 0;
  std::vector<unsigned char> ops;
  mgc_prog (g, top, 0, ops);
  if (g.gamma == 2) mgc_prog (g, top, 1, ops);
  if (ops.size () > (size_t) MGC_MAX_OPS) return 0;
  LGH_TRY (k_mg_coarse (top + 1, lev, g.cinv, g.n0, (int) ops.size (),
                        ops.data (), &g.mgc_img, &g.mgc_mu));
  *done = 1;
  return 0;
}

// the coarse correction of level l: one cycle on level l - 1, or two
// (gamma = 2) when level l - 1 is not the coarsest (exact) one
static int mg_run (DevMG &g, DevMG::Work &w, int l, const PMat &B, PMat &X,
                   int x_init);
static int
mg_coarse (DevMG &g, DevMG::Work &w, int l)
{
  int done = 0;
  LGH_TRY (mg_coarse_fused (g, w, l, &done));
  if (done) return 0;
  LGH_TRY (mg_run (g, w, l - 1, w.Bl[l - 1], w.Xl[l - 1], 0));
  if (g.gamma == 2 && l - 1 >= 1)
    LGH_TRY (mg_run (g, w, l - 1, w.Bl[l - 1], w.Xl[l - 1], 1));
  return 0;
}

// zsolve_impl.h:406-485 (X overwritten; zero initial guess).  x_init: X
// holds an initial guess and the pre-smoother starts from B - A X (the
// W-cycle's second coarse correction)
static int
mg_run (DevMG &g, DevMG::Work &w, int l, const PMat &B, PMat &X, int x_init)
{
  if (l == 0) return k_dense_apply (g.cinv, g.n0, B, X);
  if (fused_level (g, l, B.W)) {
    // degree-2 smoother as one pass (k_smooth2x), explicit residuals
    ChebCoef cc (g.slo[l], g.shi[l]);
    double a1, b1, a2, b2;
    cc.step (1, &a1, &b1);
    cc.step (2, &a2, &b2);
    const int fine = (l == g.nl - 1);
    prof_set_fine_tag (fine);
    if (x_init) {
      LGH_TRY (k_resid (*g.A[l], B, X, w.R1[l]));
      LGH_TRY (k_smooth2x (*g.A[l], b1, a2, b2, w.R1[l], X, 1));     // pre
    }
    else LGH_TRY (k_smooth2x (*g.A[l], b1, a2, b2, B, X, 0));        // pre
    LGH_TRY (k_resid (*g.A[l], B, X, w.R1[l]));
    prof_set_fine_tag (0);
    LGH_TRY (k_spmm (*g.PT[l], 1.0, w.R1[l], 0.0, w.Bl[l - 1]));
    LGH_TRY (mg_coarse (g, w, l));
    LGH_TRY (k_spmm (*g.P[l], 1.0, w.Xl[l - 1], 1.0, X));
    prof_set_fine_tag (fine);
    LGH_TRY (k_resid (*g.A[l], B, X, w.R1[l]));
    LGH_TRY (k_smooth2x (*g.A[l], b1, a2, b2, w.R1[l], X, 1));       // post
    prof_set_fine_tag (0);
    return 0;
  }
  // pre-smoothing; the residual B - A X it leaves (the Chebyshev
  // recurrence keeps it) is restricted directly — the reference forms the
  // same residual again with one more product (zsolve_impl.h:464-468)
  const PMat *R = nullptr;
  LGH_TRY (mg_smooth (g, w, l, B, X, x_init ? 0 : 1, &R));
  LGH_TRY (k_spmm (*g.PT[l], 1.0, *R, 0.0, w.Bl[l - 1]));
  LGH_TRY (mg_coarse (g, w, l));
  LGH_TRY (k_spmm (*g.P[l], 1.0, w.Xl[l - 1], 1.0, X));
  LGH_TRY (mg_smooth (g, w, l, B, X, 0, nullptr));               // post
  return 0;
}

int
mg_vcycle (DevMG &g, const PMat &B, PMat &X)
{
  LGH_CHECK (B.prec == g.prec && X.prec == g.prec, 8,
             "mg_vcycle: precision mismatch");
  DevMG::Work *w = nullptr;
  LGH_TRY (mg_work (g, B, &w));
  return mg_run (g, *w, g.nl - 1, B, X, 0);
}

ChebWork::~ChebWork ()
{
  pmat_free (R[0]); pmat_free (R[1]); pmat_free (D[0]); pmat_free (D[1]);
  pmat_free (Z); pmat_free (Rpc); pmat_free (Zpc); pmat_free (Bt);
  pmat_free (Xt);
  delete tail;
  delete tail8;
}

// whole panels and the narrow last panel of B (see zsolve.hpp): calls
// run (B', X', work) on the whole panels in place and on a 16-wide copy of
// the last panel's columns; false when B has no such panel
template <typename WK, typename F>
static int
split_tail (const PMat &B, PMat &X, WK &w, bool *done, F &&run)
{
  *done = false;
  const int full = B.ncols / 64, tail = B.ncols - 64 * full;
  if (B.W != 64 || tail == 0 || tail > 48 || !tail_split ()) return 0;
  LGH_CHECK (X.W == 64 && X.ncols == B.ncols && X.prec == B.prec, 8,
             "Z-solve: X shape differs from B");
  *done = true;
  const int tw = tail <= 8 ? 8 : 16;
  WK *&tw_work = (tw == 8) ? w.tail8 : w.tail;
  if (!tw_work) tw_work = new WK ();
  if (!w.Bt.p || w.Bt.nrows != B.nrows || w.Bt.prec != B.prec) {
    pmat_free (w.Bt); pmat_free (w.Xt);
    LGH_TRY (pmat_alloc (w.Bt, B.nrows, 48, 16, B.prec));
    LGH_TRY (pmat_alloc (w.Xt, B.nrows, 48, 16, B.prec));
  }
  // the tail's columns as one 8-wide panel or ceil(tail / 16) 16-wide ones
  // (a prefix of the 48-column buffers); Bt's columns past `tail` are zero
  // padding (so are Xt's after the solve)
  PMat Bt = w.Bt, Xt = w.Xt;
  Bt.ncols = Xt.ncols = tail;
  Bt.W = Xt.W = tw;
  LGH_TRY (dzero (Bt.p, Bt.bytes ()));
  LGH_TRY (k_copy_cols (B, 64 * full, Bt, 0, tail));
  // the narrow panel's solve runs on the auxiliary stream, overlapping the
  // whole panels' (their coarse levels are latency-bound and leave the
  // device idle); it touches only its own buffers (tail work, the V-cycle's
  // narrow-panel work slot, Bt, Xt)
  const hipStream_t main = stream ();
  hipStream_t aux = (full > 0) ? aux_stream () : nullptr;
  hipEvent_t ready = nullptr, done_ev = nullptr;
  if (aux && (hipEventCreateWithFlags (&ready, hipEventDisableTiming) != hipSuccess
              || hipEventCreateWithFlags (&done_ev, hipEventDisableTiming) != hipSuccess
              || hipEventRecord (ready, main) != hipSuccess
              || hipStreamWaitEvent (aux, ready, 0) != hipSuccess))
    aux = nullptr;                       // fall back to one stream
  int rc = 0;
  if (aux) {
    StreamScope on_aux (aux);
    rc = run (Bt, Xt, *tw_work);
    if (!rc && hipEventRecord (done_ev, aux) != hipSuccess)
      rc = hip_fail (hipGetLastError (), "hipEventRecord", __FILE__, __LINE__);
  }
  else rc = run (Bt, Xt, *tw_work);       // on the main stream, in order
  if (!rc && full > 0) {                 // the whole panels, in place
    PMat Bf = B, Xf = X;
    Bf.ncols = Xf.ncols = 64 * full;
    rc = run (Bf, Xf, w);
  }
  if (aux) {
    if (!rc && hipStreamWaitEvent (main, done_ev, 0) != hipSuccess)
      rc = hip_fail (hipGetLastError (), "hipStreamWaitEvent", __FILE__, __LINE__);
    if (rc) (void) hipStreamSynchronize (aux);
  }
  // the waits just enqueued may still be pending: the events are destroyed
  // after the next synchronization of this thread's stream
  event_retire (ready);
  event_retire (done_ev);
  LGH_TRY (rc);
  LGH_TRY (dzero (X.panel (full), (size_t) X.nrows * 64 * prec_bytes (X.prec)));
  return k_copy_cols (Xt, 0, X, 64 * full, tail);
}

int
ChebWork::ensure (const PMat &like)
{
  if (W == like.W && np == like.npanels () && prec == like.prec
      && nrows == like.nrows) return 0;
  pmat_free (R[0]); pmat_free (R[1]); pmat_free (D[0]); pmat_free (D[1]);
  pmat_free (Z); pmat_free (Rpc); pmat_free (Zpc);
  const int nc = like.npanels () * like.W;
  LGH_TRY (pmat_alloc (R[0], like.nrows, nc, like.W, like.prec));
  LGH_TRY (pmat_alloc (R[1], like.nrows, nc, like.W, like.prec));
  LGH_TRY (pmat_alloc (D[0], like.nrows, nc, like.W, like.prec));
  LGH_TRY (pmat_alloc (D[1], like.nrows, nc, like.W, like.prec));
  LGH_TRY (pmat_alloc (Z, like.nrows, nc, like.W, like.prec));
  W = like.W; np = like.npanels (); prec = like.prec; nrows = like.nrows;
  return 0;
}

static int
cheb_run (const DevCSR &A, DevMG *mg, double lmin, double lmax, int nit,
          const PMat &B, PMat &X, ChebWork &w)
{
  LGH_TRY (w.ensure (B));
  // a V-cycle in a lower storage precision (BF16 around FP32 vectors) reads
  // and writes its own copies: Rpc (written by each step) and Zpc
  const bool low = mg && mg->prec != B.prec;
  if (low) {
    LGH_CHECK ((B.W == 64 || B.W == 16 || B.W == 8) && B.prec == F32
               && mg->prec == BF16, 8,
               "cheb_solve: unsupported preconditioner precision");
    if (!w.Rpc.p) {
      const int nc = B.npanels () * B.W;
      LGH_TRY (pmat_alloc (w.Rpc, B.nrows, nc, B.W, mg->prec));
      LGH_TRY (pmat_alloc (w.Zpc, B.nrows, nc, B.W, mg->prec));
    }
    LGH_TRY (k_copy_cols (B, 0, w.Rpc, 0, B.npanels () * B.W));
  }
  ChebCoef cc (lmin, lmax);
  const PMat *Rsrc = &B;
  int ri = 0, di = 0;
  for (int k = 1; k <= nit; k++) {
    double a, b;
    cc.step (k, &a, &b);
    const PMat *Dcur = (k == 1) ? nullptr : &w.D[di ^ 1];
    if (low) {
      LGH_TRY (mg_vcycle (*mg, w.Rpc, w.Zpc));
      LGH_TRY (k_cheb_step (A, 0, &w.Zpc, a, b, Dcur, *Rsrc, X, k > 1,
                            w.D[di], w.R[ri], &w.Rpc));
    }
    else if (mg) {
      LGH_TRY (mg_vcycle (*mg, *Rsrc, w.Z));
      LGH_TRY (k_cheb_step (A, 0, &w.Z, a, b, Dcur, *Rsrc, X, k > 1, w.D[di],
                            w.R[ri]));
    }
    else
      LGH_TRY (k_cheb_step (A, 1, nullptr, a, b, Dcur, *Rsrc, X, k > 1,
                            w.D[di], w.R[ri]));
    Rsrc = &w.R[ri];
    ri ^= 1;
    di ^= 1;
  }
  return 0;
}

int
cheb_solve (const DevCSR &A, DevMG *mg, double lmin, double lmax, int nit,
            const PMat &B, PMat &X, ChebWork &w)
{
  LGH_CHECK (nit >= 1, 8, "cheb_solve: nit < 1");
  bool done = false;
  LGH_TRY (split_tail (B, X, w, &done,
                       [&] (const PMat &Bp, PMat &Xp, ChebWork &wp) {
                         return cheb_run (A, mg, lmin, lmax, nit, Bp, Xp, wp);
                       }));
  return done ? 0 : cheb_run (A, mg, lmin, lmax, nit, B, X, w);
}

ChebMixedWork::~ChebMixedWork ()
{
  pmat_free (R[0]); pmat_free (R[1]); pmat_free (D[0]); pmat_free (D[1]);
  pmat_free (Rpc); pmat_free (Zpc); pmat_free (Bt); pmat_free (Xt);
  pmat_free (R32[0]); pmat_free (R32[1]); pmat_free (X32);
  delete tail;
  delete tail8;
}

int
ChebMixedWork::ensure (const PMat &like, int pc)
{
  if (W == like.W && np == like.npanels () && nrows == like.nrows
      && pprec == pc) return 0;
  pmat_free (R[0]); pmat_free (R[1]); pmat_free (D[0]); pmat_free (D[1]);
  pmat_free (Rpc); pmat_free (Zpc);
  pmat_free (R32[0]); pmat_free (R32[1]); pmat_free (X32);
  const int nc = like.npanels () * like.W;
  LGH_TRY (pmat_alloc (D[0], like.nrows, nc, like.W, F32));
  LGH_TRY (pmat_alloc (D[1], like.nrows, nc, like.W, F32));
  LGH_TRY (pmat_alloc (Rpc, like.nrows, nc, like.W, pc));
  LGH_TRY (pmat_alloc (Zpc, like.nrows, nc, like.W, pc));
  LGH_TRY (pmat_alloc (R32[0], like.nrows, nc, like.W, F32));
  LGH_TRY (pmat_alloc (R32[1], like.nrows, nc, like.W, F32));
  LGH_TRY (pmat_alloc (X32, like.nrows, nc, like.W, F32));
  W = like.W;
  np = like.npanels ();
  nrows = like.nrows;
  pprec = pc;
  return 0;
}

int
ChebMixedWork::ensure64 ()
{
  if (R[0].p && R[0].W == W && R[0].npanels () == np && R[0].nrows == nrows)
    return 0;
  pmat_free (R[0]); pmat_free (R[1]);
  LGH_TRY (pmat_alloc (R[0], nrows, np * W, W, F64));
  return pmat_alloc (R[1], nrows, np * W, W, F64);
}

static int
cheb_mixed_run (const DevCSR &A64, DevMG &mg, double lmin, double lmax,
                int nit, const PMat &B, PMat &X, ChebMixedWork &w,
                const double *xscale)
{
  LGH_TRY (w.ensure (B, mg.prec));
  ChebCoef cc (lmin, lmax);
  LGH_TRY (k_copy_cols (B, 0, w.Rpc, 0, B.npanels () * B.W));
  // the early phase: steps 1..k_sw with FP32 R and X (k_sw even, so that X
  // holds every direction when it ends).  Half the steps: a longer phase
  // (ending when the Chebyshev bound has shrunk the error by 2^-25, 10 of 16
  // steps here) moved the smallest kept modes of a 16^3 build by 1e-7.
  // Then the residual replacement: R = B - A X formed in FP64 from the
  // FP32 X (the first X update after it reads that X and writes X in
  // FP64); the remaining steps keep X in FP64 and R in FP32.  R's rounding
  // there is relative to the residual the early phase left (about the
  // Chebyshev bound at k_sw times B), so the drift it adds to the
  // recurrence stays far below the solve's tolerance (1024^2 anisotropic
  // PSF build, 16 steps: spectrum and solves still within 1.4e-11 and
  // 3e-11 of FP64).  X stays FP64:
  // an FP32 correction accumulated after the replacement puts its rounding
  // straight into the solution (a 16^3 build extended twice moved 5x more).
  const int k_sw = (nit >= 8) ? 2 * (nit / 4) : 0;
  if (!k_sw) LGH_TRY (w.ensure64 ());
  const PMat *Rsrc = &B;
  int ri = 0, di = 0;
  bool x_set = false;               // X (FP64) written since the replacement
  for (int k = 1; k <= nit; k++) {
    double a, b;
    cc.step (k, &a, &b);
    const PMat *Dcur = (k == 1) ? nullptr : &w.D[di ^ 1];
    // X takes the directions two at a time: odd steps before the last leave
    // X alone, the next step adds both (same sums in the same order)
    int xm;
    if (k % 2 == 1 && k < nit) xm = MIX_X_SKIP;
    else if (k % 2 == 0) xm = (k == 2) ? MIX_X_SET2 : MIX_X_ADD2;
    else xm = (k == 1) ? MIX_X_SET : MIX_X_ADD;
    const bool early = k <= k_sw;
    PMat &Xk = early ? w.X32 : X;
    PMat &Rn = k_sw ? w.R32[ri] : w.R[ri];
    // the first X update after the replacement reads the early phase's X32
    // (widened exactly) instead of an FP64 copy of it
    const bool first_x = k_sw && !x_set && !early && xm != MIX_X_SKIP;
    LGH_TRY (mg_vcycle (mg, w.Rpc, w.Zpc));
    LGH_TRY (k_cheb_step_mixed (A64, w.Zpc, a, b, Dcur, *Rsrc, Xk, xm,
                                w.D[di], Rn, w.Rpc, first_x ? &w.X32 : nullptr,
                                k == nit ? xscale : nullptr));
    if (first_x) x_set = true;
    Rsrc = &Rn;
    ri ^= 1;
    di ^= 1;
    if (k == k_sw) {
      // residual replacement: R = B - A X32 in FP64 (stored in FP32), and
      // the next preconditioner input from it (one pass); X in FP64 is
      // first written by the next update that adds to it
      LGH_TRY (k_mix_switch (A64, B, w.X32, w.R32[ri], w.Rpc));
      Rsrc = &w.R32[ri];
      ri ^= 1;
    }
  }
  return 0;
}

int
cheb_solve_mixed (const DevCSR &A64, DevMG &mg, double lmin, double lmax,
                  int nit, const PMat &B, PMat &X, ChebMixedWork &w,
                  const double *xscale)
{
  LGH_CHECK (nit >= 1 && B.W == 64 && X.W == 64 && B.prec == F64
             && X.prec == F64 && B.ncols == X.ncols
             && (mg.prec == F32 || mg.prec == BF16), 8,
             "cheb_solve_mixed: shape/precision");
  bool done = false;
  LGH_TRY (split_tail (B, X, w, &done,
                       [&] (const PMat &Bp, PMat &Xp, ChebMixedWork &wp) {
                         return cheb_mixed_run (A64, mg, lmin, lmax, nit, Bp,
                                                Xp, wp, xscale);
                       }));
  return done ? 0 : cheb_mixed_run (A64, mg, lmin, lmax, nit, B, X, w, xscale);
}

int
pcg_solve (const DevCSR &A, DevMG *mg, const double *dinv64, const PMat &B,
           PMat &X, double rtol, int maxit, PcgInfo *info)
{
  const int nc = B.npanels () * B.W;
  PMat R, Z, P, Q;
  int rc = 0;
  std::vector<double> dp (nc), pq (nc), dn (nc), zn (nc), z0 (nc);
  std::vector<double> al (nc), be (nc);
  std::vector<int> done (nc, 0);
  double *dev = (double *) dmalloc (sizeof (double) * 4 * (size_t) nc);
  if (!dev) return 5;
  double *d_a = dev, *d_b = dev + nc, *d_t1 = dev + 2 * nc, *d_t2 = dev + 3 * nc;
  auto precond = [&] (const PMat &r, PMat &z) -> int {
    if (mg) return mg_vcycle (*mg, r, z);
    return k_diag_scale (r, dinv64, z);
  };
  if (info) *info = PcgInfo ();
#define PCG_TRY(e) do { rc = (e); if (rc) goto out; } while (0)
  PCG_TRY (pmat_alloc (R, B.nrows, nc, B.W, B.prec));
  PCG_TRY (pmat_alloc (Z, B.nrows, nc, B.W, B.prec));
  PCG_TRY (pmat_alloc (P, B.nrows, nc, B.W, B.prec));
  PCG_TRY (pmat_alloc (Q, B.nrows, nc, B.W, B.prec));
  PCG_TRY (dzero (X.p, X.bytes ()));
  PCG_TRY (dcopy (R.p, B.p, B.bytes ()));
  PCG_TRY (precond (R, Z));
  PCG_TRY (k_coldot (R, Z, d_t1));
  PCG_TRY (k_coldot (Z, Z, d_t2));
  PCG_TRY (d2h (dp.data (), d_t1, sizeof (double) * nc));
  PCG_TRY (d2h (z0.data (), d_t2, sizeof (double) * nc));
  for (int j = 0; j < nc; j++) {
    z0[j] = std::sqrt (z0[j]);
    if (!(z0[j] > 0.0)) done[j] = 1;
  }
  PCG_TRY (dcopy (P.p, Z.p, Z.bytes ()));
  {
    int it = 0, all = 1;
    for (int j = 0; j < nc; j++) all &= done[j];
    while (!all && it < maxit) {
      PCG_TRY (k_spmm (A, 1.0, P, 0.0, Q));
      PCG_TRY (k_coldot (P, Q, d_t1));
      PCG_TRY (d2h (pq.data (), d_t1, sizeof (double) * nc));
      for (int j = 0; j < nc; j++) al[j] = done[j] ? 0.0 : dp[j] / pq[j];
      PCG_TRY (h2d (d_a, al.data (), sizeof (double) * nc));
      PCG_TRY (k_cg_xr (d_a, P, Q, X, R));
      PCG_TRY (precond (R, Z));
      PCG_TRY (k_coldot (R, Z, d_t1));
      PCG_TRY (k_coldot (Z, Z, d_t2));
      PCG_TRY (d2h (dn.data (), d_t1, sizeof (double) * nc));
      PCG_TRY (d2h (zn.data (), d_t2, sizeof (double) * nc));
      for (int j = 0; j < nc; j++) be[j] = done[j] ? 0.0 : dn[j] / dp[j];
      if (info && !done[0]) { info->alpha.push_back (al[0]); info->beta.push_back (be[0]); }
      it++;
      all = 1;
      for (int j = 0; j < nc; j++) {
        if (!done[j] && std::sqrt (zn[j]) <= rtol * z0[j]) done[j] = 1;
        all &= done[j];
        if (!done[j]) dp[j] = dn[j];
      }
      if (all) break;
      PCG_TRY (h2d (d_b, be.data (), sizeof (double) * nc));
      PCG_TRY (k_cg_p (d_b, Z, P));
    }
    if (info) { info->its = it; info->converged = all; }
  }
out:
#undef PCG_TRY
  pmat_free (R); pmat_free (Z); pmat_free (P); pmat_free (Q);
  dfree (dev);
  return rc;
}

}  // namespace lgh
