// capi.cpp — the lgh_hip_* / lgh_amg_* C ABI over the internal device
// layer: host column-major blocks are staged into panel matrices, the
// kernel runs on the engine stream, results are copied back.
#include <lgh_hip.h>

#include "engine.hpp"
#include "randn.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

using namespace lgh;

namespace {

// host column-major block -> device panel matrix (prec, W)
int
stage_in (const double *X, int ld, int n, int nc, int W, int prec, PMat &out)
{
  PMat H;
  std::vector<double> h ((size_t) n * nc);
  for (int j = 0; j < nc; j++) memcpy (&h[(size_t) j * n], X + (size_t) j * ld, sizeof (double) * n);
  LGH_TRY (pmat_alloc (H, n, nc, 1, F64));
  int rc = h2d (H.p, h.data (), sizeof (double) * h.size ());
  if (!rc) rc = pmat_alloc (out, n, nc, W, prec);
  if (!rc) rc = k_copy_cols (H, 0, out, 0, nc);
  if (!rc) rc = dev_sync ();
  pmat_free (H);
  return rc;
}

int
stage_out (const PMat &X, int nc, double *Y, int ld)
{
  PMat H;
  const int n = X.nrows;
  std::vector<double> h ((size_t) n * nc);
  LGH_TRY (pmat_alloc (H, n, nc, 1, F64));
  int rc = k_copy_cols (X, 0, H, 0, nc);
  if (!rc) rc = d2h (h.data (), H.p, sizeof (double) * h.size ());
  pmat_free (H);
  LGH_TRY (rc);
  for (int j = 0; j < nc; j++) memcpy (Y + (size_t) j * ld, &h[(size_t) j * n], sizeof (double) * n);
  return 0;
}

}  // namespace

struct lgh_amg {
  AmgHierarchy h;
  DevMG *mg[3] = { nullptr, nullptr, nullptr };   // per precision
  ~lgh_amg () { for (auto *g : mg) delete g; }
};

extern "C" {

int lgh_hip_init (void) { return dev_init (); }
const char *lgh_hip_last_error (void) { return lgh::last_error (); }
int lgh_hip_sync (void) { return dev_sync (); }
int64_t lgh_hip_bytes_in_use (void) { return (int64_t) dbytes_in_use (); }
void lgh_hip_trim (void) { dtrim (); }

void lgh_hip_prof_enable (int on) { prof_enable (on); }

void *
lgh_hip_dev_alloc (int64_t bytes)
{
  void *p = nullptr;
  return dev_block_alloc (bytes > 0 ? (size_t) bytes : 0, &p) ? nullptr : p;
}

void lgh_hip_dev_free (void *p) { dev_block_free (p); }

int
lgh_hip_dev_put (void *dst, const void *src, int64_t bytes)
{
  LGH_TRY (dev_init ());
  LGH_CHECK (bytes >= 0 && device_range_ok (dst, (size_t) bytes), 8,
             "lgh_hip_dev_put: dst must be device memory holding bytes");
  return h2d (dst, src, (size_t) bytes);
}

int
lgh_hip_dev_get (void *dst, const void *src, int64_t bytes)
{
  LGH_TRY (dev_init ());
  LGH_CHECK (bytes >= 0 && device_range_ok (src, (size_t) bytes), 8,
             "lgh_hip_dev_get: src must be device memory holding bytes");
  return d2h (dst, src, (size_t) bytes);
}

int
lgh_hip_copy_gbps (int64_t bytes, int reps, double *gbps)
{
  LGH_TRY (dev_init ());
  return copy_gbps ((size_t) bytes, reps, gbps);
}

int
lgh_hip_prof_collect (int kind, int np, int prec, int64_t *count, double *ms,
                      double *bytes, double *flops)
{
  LGH_CHECK (kind >= 0 && kind < PK_N, 8, "lgh_hip_prof_collect: bad kind");
  long c = 0;
  LGH_TRY (prof_collect (kind, np, prec, &c, ms, bytes, flops));
  *count = c;
  return 0;
}

double
lgh_hip_randn_at_host (uint64_t seed, int64_t i, int64_t j)
{
  return lgh::randn_at (seed, i, j);
}

int
lgh_hip_randn_host_block (uint64_t seed, int64_t row0, int64_t col0, int nr,
                          int nc, int prec, double *vals, int ld,
                          unsigned char *flags, int ldf)
{
  LGH_CHECK (nr >= 0 && nc >= 0 && ld >= std::max (nr, 1)
             && ldf >= std::max (nr, 1) && (prec == 0 || prec == 1), 63,
             "lgh_hip_randn_host_block: bad arguments");
  auto col = [&] (int j) {
    for (int i = 0; i < nr; i++) {
      bool f;
      double u1, q;
      const double v = prec == 0
          ? lgh::randn_at_flag<double> (seed, row0 + i, col0 + j, &f, &u1, &q)
          : (double) lgh::randn_at_flag<float> (seed, row0 + i, col0 + j, &f, &u1, &q);
      vals[(size_t) j * ld + i] = v;
      flags[(size_t) j * ldf + i] = f ? 1 : 0;
    }
  };
  const int nt = (int) std::max (1u, std::min (8u, std::thread::hardware_concurrency ()));
  std::atomic<int> next (0);
  auto worker = [&] () { for (int j; (j = next++) < nc;) col (j); };
  std::vector<std::thread> th;
  for (int t = 1; t < std::min (nt, std::max (nc, 1)); t++) th.emplace_back (worker);
  worker ();
  for (auto &t : th) t.join ();
  return 0;
}

int
lgh_hip_randn_fill (uint64_t seed, int64_t row0, int64_t col0, int nr, int nc,
                    int W, int prec, double *out, int ld)
{
  LGH_TRY (dev_init ());
  PMat X;
  LGH_TRY (pmat_alloc (X, nr, nc, W, prec));
  int rc = k_randn (X, seed, row0, col0);
  if (!rc) rc = stage_out (X, nc, out, ld);
  pmat_free (X);
  return rc;
}

int
lgh_hip_randn_fill_noalloc (uint64_t seed, int64_t row0, int64_t col0, int nr,
                            int nc, int prec, void *out, int ld)
{
  return randn_fill_noalloc (seed, row0, col0, nr, nc, prec, out, ld);
}

static int
spmm_host (int n, int m, const int *rp, const int *ci, const double *v,
           int prec_vals, int prec_vec, int W, int ncols, const double *X,
           int ldx, double *Y, int ldy, int grouped)
{
  LGH_TRY (dev_init ());
  DevCSR *A = nullptr;
  LGH_TRY (csr_upload (n, m, rp, ci, v, prec_vals, 0, &A));
  std::unique_ptr<DevCSR> hold (A);
  if (grouped) LGH_TRY (csr_group_rows (*A, rp, ci, v));
  PMat dx, dy;
  int rc = stage_in (X, ldx, m, ncols, W, prec_vec, dx);
  if (!rc) rc = pmat_alloc (dy, n, ncols, W, prec_vec);
  if (!rc) rc = k_spmm (*A, 1.0, dx, 0.0, dy);
  if (!rc) rc = stage_out (dy, ncols, Y, ldy);
  pmat_free (dx); pmat_free (dy);
  return rc;
}

void
lgh_hip_spmm_group_config (int rg, int ppw)
{
  spmm_group_force (rg, ppw);
}

void
lgh_hip_spmm_tile_config (int cand)
{
  spmm_tile_force (cand);
}

void
lgh_hip_mg_coarse_config (int max_rows, int max_nnz)
{
  mg_coarse_config (max_rows, max_nnz);
}

int
lgh_hip_spmm_group_report (char *buf, int len)
{
  const std::string r = spmm_group_report ();
  if (buf && len > 0) {
    const size_t k = std::min ((size_t) len - 1, r.size ());
    memcpy (buf, r.data (), k);
    buf[k] = 0;
  }
  return (int) r.size ();
}

int
lgh_hip_spmm (int n, int m, const int *rp, const int *ci, const double *v,
              int prec_vals, int prec_vec, int W, int ncols, const double *X,
              int ldx, double *Y, int ldy)
{
  return spmm_host (n, m, rp, ci, v, prec_vals, prec_vec, W, ncols, X, ldx, Y,
                    ldy, 0);
}

int
lgh_hip_spmm_grouped (int n, int m, const int *rp, const int *ci,
                      const double *v, int prec_vals, int prec_vec, int W,
                      int ncols, const double *X, int ldx, double *Y, int ldy)
{
  return spmm_host (n, m, rp, ci, v, prec_vals, prec_vec, W, ncols, X, ldx, Y,
                    ldy, 1);
}

int
lgh_hip_cheb_jacobi (int n, const int *rp, const int *ci, const double *v,
                     int prec, int W, double lmin, double lmax, int nit,
                     int ncols, const double *B, int ldb, double *X, int ldx)
{
  LGH_TRY (dev_init ());
  DevCSR *A = nullptr;
  LGH_TRY (csr_upload (n, n, rp, ci, v, prec, 1, &A));
  std::unique_ptr<DevCSR> hold (A);
  PMat db, dx;
  ChebWork w;
  int rc = stage_in (B, ldb, n, ncols, W, prec, db);
  if (!rc) rc = pmat_alloc (dx, n, ncols, W, prec);
  if (!rc) rc = cheb_solve (*A, nullptr, lmin, lmax, nit, db, dx, w);
  if (!rc) rc = stage_out (dx, ncols, X, ldx);
  pmat_free (db); pmat_free (dx);
  return rc;
}

int
lgh_hip_gram (int prec, int n, int a, const double *X, int ldx, int b,
              const double *Y, int ldy, double *G, int ldg)
{
  LGH_TRY (dev_init ());
  PMat dx, dy;
  const int Pa = (a + 63) / 64, Pb = (b + 63) / 64;
  const int ld = Pa * 64;
  double *dG = (double *) dmalloc (sizeof (double) * (size_t) ld * Pb * 64);
  if (!dG) return 5;
  std::vector<double> h ((size_t) ld * Pb * 64);
  int rc = stage_in (X, ldx, n, a, 64, prec, dx);
  if (!rc) rc = stage_in (Y, ldy, n, b, 64, prec, dy);
  if (!rc) rc = k_gram (dx, 0, Pa, dy, 0, Pb, dG, ld, 0);
  if (!rc) rc = d2h (h.data (), dG, sizeof (double) * h.size ());
  pmat_free (dx); pmat_free (dy); dfree (dG);
  LGH_TRY (rc);
  for (int j = 0; j < b; j++)
    for (int i = 0; i < a; i++) G[i + (size_t) j * ldg] = h[i + (size_t) j * ld];
  return 0;
}

int
lgh_hip_gemm_nn (int prec, int n, int a, const double *X, int ldx, int b,
                 const double *C, int ldc, double *Y, int ldy)
{
  LGH_TRY (dev_init ());
  PMat dx, dy;
  const int Pa = (a + 63) / 64, Pb = (b + 63) / 64;
  const int ld = Pa * 64;
  std::vector<double> hc ((size_t) ld * Pb * 64, 0.0);
  for (int j = 0; j < b; j++)
    for (int i = 0; i < a; i++) hc[i + (size_t) j * ld] = C[i + (size_t) j * ldc];
  // a C whose 64 x 64 blocks below the block diagonal are all zero takes
  // the triangular product (the form CholQR uses)
  int upper = 1;
  for (int j = 0; j < Pb * 64 && upper; j++)
    for (int i = (j / 64 + 1) * 64; i < ld; i++)
      if (hc[i + (size_t) j * ld] != 0.0) { upper = 0; break; }
  double *dC = (double *) dmalloc (sizeof (double) * hc.size ());
  if (!dC) return 5;
  int rc = h2d (dC, hc.data (), sizeof (double) * hc.size ());
  if (!rc) rc = stage_in (X, ldx, n, a, 64, prec, dx);
  if (!rc) rc = pmat_alloc (dy, n, b, 64, prec);
  if (!rc) rc = k_gemm_nn (dx, 0, Pa, dC, ld, 1.0, 0.0, dy, 0, Pb, upper);
  if (!rc) rc = stage_out (dy, b, Y, ldy);
  pmat_free (dx); pmat_free (dy); dfree (dC);
  return rc;
}

int
lgh_amg_build (int n, const int *rp, const int *ci, const double *v,
               int smax_krylov, double smoother_top, int nu_force,
               lgh_amg_t **out)
{
  lgh_amg_t *a = new lgh_amg_t ();
  HostCSR A;
  A.n = n; A.m = n;
  A.rp.assign (rp, rp + n + 1);
  A.ci.assign (ci, ci + rp[n]);
  A.v.assign (v, v + rp[n]);
  AmgOpts o;
  o.smax_krylov = smax_krylov;
  o.smoother_top = smoother_top;
  o.nu_force = nu_force;
  if (amg_setup (A, o, a->h)) {
    delete a;
    set_error ("AMG setup failed (coarsening stalled above 2000 unknowns)");
    return 13;
  }
  *out = a;
  return 0;
}

void lgh_amg_free (lgh_amg_t *h) { delete h; }
int lgh_amg_levels (const lgh_amg_t *h) { return (int) h->h.lev.size (); }

int
lgh_amg_level (const lgh_amg_t *h, int l, int *n, int *nc, const int **rp,
               const int **ci, const double **v, const int **prp,
               const int **pci, const double **pv, double *slo, double *shi,
               int *nu)
{
  LGH_CHECK (l >= 0 && l < (int) h->h.lev.size (), 8, "lgh_amg_level: range");
  const AmgLevel &lv = h->h.lev[l];
  *n = lv.A.n;
  *nc = l > 0 ? lv.P.m : 0;
  *rp = lv.A.rp.data (); *ci = lv.A.ci.data (); *v = lv.A.v.data ();
  *prp = l > 0 ? lv.P.rp.data () : nullptr;
  *pci = l > 0 ? lv.P.ci.data () : nullptr;
  *pv = l > 0 ? lv.P.v.data () : nullptr;
  *slo = lv.slo; *shi = lv.shi; *nu = lv.nu;
  return 0;
}

int
lgh_amg_set_cycle (lgh_amg_t *h, int gamma)
{
  LGH_CHECK (gamma == 1 || gamma == 2, 8, "lgh_amg_set_cycle: gamma must be 1 or 2");
  h->h.gamma = gamma;
  for (auto *g : h->mg) if (g) g->gamma = gamma;
  return 0;
}

int
lgh_amg_vcycle (const lgh_amg_t *h, int prec, int W, int ncols,
                const double *B, int ldb, double *X, int ldx)
{
  LGH_TRY (dev_init ());
  LGH_CHECK (prec >= 0 && prec <= 2 && (prec != 2 || W == 64 || W == 16 || W == 8), 8,
             "lgh_amg_vcycle: prec must be 0, 1 or 2 (BF16 needs W = 64, 16 or 8)");
  lgh_amg_t *m = (lgh_amg_t *) h;
  if (!m->mg[prec]) LGH_TRY (mg_upload (m->h, prec, &m->mg[prec]));
  const int n = m->h.lev.back ().A.n;
  PMat db, dx;
  int rc = stage_in (B, ldb, n, ncols, W, prec, db);
  if (!rc) rc = pmat_alloc (dx, n, ncols, W, prec);
  if (!rc) rc = mg_vcycle (*m->mg[prec], db, dx);
  if (!rc) rc = stage_out (dx, ncols, X, ldx);
  pmat_free (db); pmat_free (dx);
  return rc;
}

}  // extern "C"
