// amg.cpp — host smoothed-aggregation setup (strength graph, three-pass
// greedy aggregation, piecewise-constant tentative prolongator, one damped
// Jacobi smoothing step, Galerkin coarse operators) plus the per-level
// Chebyshev smoother bounds of zsolve_impl.h:278-342.
#include "amg.hpp"
#include "randn.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace lgh {

void
host_spmv (const HostCSR &A, const double *x, double *y)
{
  for (int i = 0; i < A.n; i++) {
    double s = 0.0;
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++) s += A.v[k] * x[A.ci[k]];
    y[i] = s;
  }
}

HostCSR
host_transpose (const HostCSR &A)
{
  HostCSR T;
  T.n = A.m; T.m = A.n;
  T.rp.assign ((size_t) T.n + 1, 0);
  for (int64_t k = 0; k < A.nnz (); k++) T.rp[A.ci[k] + 1]++;
  for (int i = 0; i < T.n; i++) T.rp[i + 1] += T.rp[i];
  T.ci.resize ((size_t) A.nnz ());
  T.v.resize ((size_t) A.nnz ());
  std::vector<int> pos (T.rp.begin (), T.rp.end () - 1);
  for (int i = 0; i < A.n; i++)
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
      int p = pos[A.ci[k]]++;
      T.ci[p] = i;
      T.v[p] = A.v[k];
    }
  return T;
}

// Gustavson row-by-row product with a dense marker; columns sorted per row
HostCSR
host_spgemm (const HostCSR &A, const HostCSR &B)
{
  HostCSR C;
  C.n = A.n; C.m = B.m;
  C.rp.assign ((size_t) C.n + 1, 0);
  std::vector<int> mark ((size_t) B.m, -1);
  std::vector<double> acc ((size_t) B.m, 0.0);
  std::vector<int> cols;
  for (int i = 0; i < A.n; i++) {
    cols.clear ();
    for (int ka = A.rp[i]; ka < A.rp[i + 1]; ka++) {
      const int r = A.ci[ka];
      const double av = A.v[ka];
      for (int kb = B.rp[r]; kb < B.rp[r + 1]; kb++) {
        const int c = B.ci[kb];
        if (mark[c] != i) { mark[c] = i; acc[c] = 0.0; cols.push_back (c); }
        acc[c] += av * B.v[kb];
      }
    }
    std::sort (cols.begin (), cols.end ());
    for (int c : cols) { C.ci.push_back (c); C.v.push_back (acc[c]); }
    C.rp[i + 1] = (int) C.ci.size ();
  }
  return C;
}

// number of eigenvalues of the tridiagonal (d, e) below x (Sturm count)
static int
sturm_count (int n, const double *d, const double *e, double x)
{
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0) cnt++;
  for (int i = 1; i < n; i++) {
    double qq = (q != 0.0) ? q : 1e-300;
    q = d[i] - x - e[i - 1] * e[i - 1] / qq;
    if (q < 0) cnt++;
  }
  return cnt;
}

void
tridiag_extremes (int n, const double *d, const double *e, double *lo,
                  double *hi)
{
  double gl = d[0], gu = d[0];
  for (int i = 0; i < n; i++) {
    double r = (i > 0 ? std::fabs (e[i - 1]) : 0.0)
               + (i + 1 < n ? std::fabs (e[i]) : 0.0);
    gl = std::min (gl, d[i] - r);
    gu = std::max (gu, d[i] + r);
  }
  // smallest: largest x with count(x) == 0 ; largest: smallest x with count == n
  double a = gl, b = gu;
  for (int it = 0; it < 200 && b - a > 1e-15 * std::max (std::fabs (a), std::fabs (b)); it++) {
    double m = 0.5 * (a + b);
    if (sturm_count (n, d, e, m) >= 1) b = m; else a = m;
  }
  *lo = 0.5 * (a + b);
  a = gl; b = gu;
  for (int it = 0; it < 200 && b - a > 1e-15 * std::max (std::fabs (a), std::fabs (b)); it++) {
    double m = 0.5 * (a + b);
    if (sturm_count (n, d, e, m) >= n) b = m; else a = m;
  }
  *hi = 0.5 * (a + b);
}

int
host_cg_lanczos (const HostCSR &A, const double *dinv, const double *b,
                 double rtol, int maxit, double *emin, double *emax)
{
  const int n = A.n;
  std::vector<double> x (n, 0.0), r (b, b + n), z (n), p (n), q (n);
  std::vector<double> al, be;
  for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
  double dp = 0.0, z0 = 0.0;
  for (int i = 0; i < n; i++) { dp += r[i] * z[i]; z0 += z[i] * z[i]; }
  z0 = std::sqrt (z0);
  p = z;
  int it = 0;
  if (z0 > 0.0) {
    while (it < maxit) {
      host_spmv (A, p.data (), q.data ());
      double pq = 0.0;
      for (int i = 0; i < n; i++) pq += p[i] * q[i];
      const double a = dp / pq;
      for (int i = 0; i < n; i++) { x[i] += a * p[i]; r[i] -= a * q[i]; }
      for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
      double dn = 0.0, zn = 0.0;
      for (int i = 0; i < n; i++) { dn += r[i] * z[i]; zn += z[i] * z[i]; }
      const double beta = dn / dp;
      al.push_back (a); be.push_back (beta);
      it++;
      if (std::sqrt (zn) <= rtol * z0) break;
      for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
      dp = dn;
    }
  }
  if (it == 0) { *emin = *emax = 1.0; return 0; }
  std::vector<double> d (it), e (it);
  for (int k = 0; k < it; k++) {
    d[k] = 1.0 / al[k] + (k > 0 ? be[k - 1] / al[k - 1] : 0.0);
    e[k] = std::sqrt (be[k]) / al[k];
  }
  tridiag_extremes (it, d.data (), e.data (), emin, emax);
  return it;
}

static void
diag_inverse (const HostCSR &A, std::vector<double> &dinv)
{
  dinv.assign (A.n, 0.0);
  for (int i = 0; i < A.n; i++)
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
      if (A.ci[k] == i) dinv[i] += A.v[k];
  for (int i = 0; i < A.n; i++) dinv[i] = 1.0 / dinv[i];
}

// three-pass greedy aggregation on the strength graph
static int
aggregate (const HostCSR &A, double theta, std::vector<int> &agg)
{
  const int n = A.n;
  std::vector<double> diag (n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
      if (A.ci[k] == i) diag[i] = A.v[k];
  auto strong = [&] (int i, int k) {
    const int j = A.ci[k];
    if (j == i || A.v[k] == 0.0) return false;
    return std::fabs (A.v[k]) >= theta * std::sqrt (std::fabs (diag[i] * diag[j]));
  };
  agg.assign (n, -1);
  int na = 0;
  for (int i = 0; i < n; i++) {                     // pass 1: root + nbhd
    if (agg[i] != -1) continue;
    bool free_nb = true, any = false;
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
      if (strong (i, k)) { any = true; if (agg[A.ci[k]] != -1) { free_nb = false; break; } }
    if (!any || !free_nb) continue;
    agg[i] = na;
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
      if (strong (i, k)) agg[A.ci[k]] = na;
    na++;
  }
  std::vector<int> snap (agg);
  for (int i = 0; i < n; i++) {                     // pass 2: join strongest
    if (agg[i] != -1) continue;
    double best = -1.0;
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
      if (strong (i, k) && snap[A.ci[k]] != -1 && std::fabs (A.v[k]) > best) {
        best = std::fabs (A.v[k]);
        agg[i] = snap[A.ci[k]];
      }
  }
  for (int i = 0; i < n; i++) {                     // pass 3: leftovers
    if (agg[i] != -1) continue;
    agg[i] = na;
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
      if (strong (i, k) && agg[A.ci[k]] == -1) agg[A.ci[k]] = na;
    na++;
  }
  return na;
}

static HostCSR
smoothed_prolongator (const HostCSR &A, const std::vector<int> &agg, int na)
{
  const int n = A.n;
  std::vector<int> size (na, 0);
  for (int i = 0; i < n; i++) size[agg[i]]++;
  std::vector<double> pt (n);
  for (int i = 0; i < n; i++) pt[i] = 1.0 / std::sqrt ((double) size[agg[i]]);
  std::vector<double> dinv;
  diag_inverse (A, dinv);
  // omega = 4/3 / rho(D^-1 A), rho from 20 steps of Jacobi-preconditioned
  // CG-Lanczos (a Krylov estimate, as GAMG's eigenvalue estimate for the
  // prolongator smoother), times 1.05.  The max-absolute-row-sum bound is
  // exact on the fine 5- and 7-point stencils but overestimates rho by up
  // to ~1.8x on the Galerkin coarse operators, which under-smooths P there
  // (preconditioned kappa 25 instead of 6 on a 64^3 grid).  The estimate
  // never exceeds the row-sum bound.
  double rs = 0.0;
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++) s += std::fabs (A.v[k]);
    rs = std::max (rs, s * std::fabs (dinv[i]));
  }
  double rho = rs;
  if (n > 1) {
    std::vector<double> rb (n);
    for (int i = 0; i < n; i++) rb[i] = randn_at_libm (0x5A3C0DEULL, i, 0);
    double emin = 0.0, emax = 0.0;
    host_cg_lanczos (A, dinv.data (), rb.data (), 1e-12, 20, &emin, &emax);
    if (emax > 0.0) rho = std::min (rs, 1.05 * emax);
  }
  const double omega = (rho > 0.0) ? (4.0 / 3.0) / rho : 0.0;
  HostCSR P;
  P.n = n; P.m = na;
  P.rp.assign ((size_t) n + 1, 0);
  std::vector<int> mark (na, -1);
  std::vector<double> acc (na, 0.0);
  std::vector<int> cols;
  for (int i = 0; i < n; i++) {
    cols.clear ();
    auto add = [&] (int c, double v) {
      if (mark[c] != i) { mark[c] = i; acc[c] = 0.0; cols.push_back (c); }
      acc[c] += v;
    };
    add (agg[i], pt[i]);
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
      const int j = A.ci[k];
      add (agg[j], -omega * dinv[i] * A.v[k] * pt[j]);
    }
    std::sort (cols.begin (), cols.end ());
    for (int c : cols)
      if (acc[c] != 0.0) { P.ci.push_back (c); P.v.push_back (acc[c]); }
    P.rp[i + 1] = (int) P.ci.size ();
  }
  return P;
}

static int
dense_inverse (const HostCSR &A, std::vector<double> &inv)
{
  const int n = A.n;
  std::vector<double> M ((size_t) n * 2 * n, 0.0);
  for (int i = 0; i < n; i++) {
    for (int k = A.rp[i]; k < A.rp[i + 1]; k++) M[(size_t) i * 2 * n + A.ci[k]] += A.v[k];
    M[(size_t) i * 2 * n + n + i] = 1.0;
  }
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs (M[(size_t) r * 2 * n + c]) > std::fabs (M[(size_t) p * 2 * n + c])) p = r;
    if (M[(size_t) p * 2 * n + c] == 0.0) return 1;
    if (p != c)
      for (int k = 0; k < 2 * n; k++) std::swap (M[(size_t) p * 2 * n + k], M[(size_t) c * 2 * n + k]);
    const double piv = 1.0 / M[(size_t) c * 2 * n + c];
    for (int k = 0; k < 2 * n; k++) M[(size_t) c * 2 * n + k] *= piv;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = M[(size_t) r * 2 * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < 2 * n; k++) M[(size_t) r * 2 * n + k] -= f * M[(size_t) c * 2 * n + k];
    }
  }
  inv.resize ((size_t) n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) inv[(size_t) i * n + j] = M[(size_t) i * 2 * n + n + j];
  return 0;
}

double
amg_cycle_cost (const AmgHierarchy &h, int gamma)
{
  // 5 us per launch at 5 TB/s, shared by the 448 columns of a 400-wide
  // sketch block; a level visit is ~7 launches (smoother, residuals,
  // transfers)
  const double launch = 5e-6 * 5e12 / 448.0;
  double v = 1.0, c = 0.0;
  for (int l = (int) h.lev.size () - 1; l >= 0; l--) {
    const AmgLevel &lv = h.lev[l];
    const double n = lv.A.n, nnz = (double) lv.A.nnz ();
    if (l == 0) {
      c += v * (n * n * 4.0 / 64.0 + launch);
      break;
    }
    // 14 BF16 vector passes (smoothers, two residuals, both transfers), A
    // read 4 times and P twice per 64-column panel
    c += v * (14.0 * 2.0 * n + (32.0 * nnz + 16.0 * (double) lv.P.nnz ()) / 64.0
              + 7.0 * launch);
    if (gamma == 2 && l - 1 >= 1) v *= 2.0;
  }
  return c;
}

int
amg_setup (const HostCSR &A0, const AmgOpts &o, AmgHierarchy &h)
{
  std::vector<HostCSR> As, Ps;    // fine to coarse
  As.push_back (A0);
  while ((int) As.size () < o.max_levels && As.back ().n > o.coarse_max) {
    const HostCSR &A = As.back ();
    std::vector<int> agg;
    const int na = aggregate (A, o.strength, agg);
    if (na >= A.n || na == 0 || na > (int) (0.9 * A.n)) break;   // stalled
    // a handful of aggregates is a poor coarse space: keep this level as
    // the coarsest one (solved exactly by its dense inverse) when it is
    // small enough for that
    if (na < 16 && A.n <= 2000) break;
    HostCSR P = smoothed_prolongator (A, agg, na);
    HostCSR AP = host_spgemm (A, P);
    HostCSR PT = host_transpose (P);
    HostCSR Ac = host_spgemm (PT, AP);
    Ps.push_back (std::move (P));
    As.push_back (std::move (Ac));
  }
  const int nl = (int) As.size ();
  if (As.back ().n > 2000) return 1;    // coarse dense solve needs n0 <= 2000
  h.lev.assign (nl, AmgLevel ());
  for (int f = 0; f < nl; f++) {        // f: fine index; l: PCMG index
    const int l = nl - 1 - f;
    AmgLevel &lv = h.lev[l];
    lv.A = As[f];
    if (l > 0) lv.P = Ps[f];
    diag_inverse (lv.A, lv.dinv);
    if (l > 0) {
      lv.nu = o.nu_force > 0 ? o.nu_force : 2;
      std::vector<double> rb (lv.A.n);
      for (int i = 0; i < lv.A.n; i++) rb[i] = randn_at_libm (20260813ULL + (uint64_t) l, i, 0);
      if (o.smax_krylov) {
        double emin, emax;
        host_cg_lanczos (lv.A, lv.dinv.data (), rb.data (), 1e-12, 40, &emin, &emax);
        lv.smax = emax;
      }
      else {
        std::vector<double> v (rb), w (lv.A.n);
        double nrm = 0.0;
        for (double x : v) nrm += x * x;
        nrm = std::sqrt (nrm);
        for (double &x : v) x /= nrm;
        for (int it = 0; it < 20; it++) {
          host_spmv (lv.A, v.data (), w.data ());
          nrm = 0.0;
          for (int i = 0; i < lv.A.n; i++) { w[i] *= lv.dinv[i]; nrm += w[i] * w[i]; }
          nrm = std::sqrt (nrm);
          for (int i = 0; i < lv.A.n; i++) v[i] = w[i] / nrm;
        }
        lv.smax = nrm;
      }
      lv.slo = 0.1 * lv.smax;
      lv.shi = o.smoother_top * lv.smax;
    }
  }
  return dense_inverse (h.lev[0].A, h.coarse_inv) ? 2 : 0;
}

}  // namespace lgh
