/* lgh_oracle.c — FP64 CPU restatement of the reference GLR path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity reference and the timed CPU baseline).
 * The product library never links or calls it.  Each function cites the
 * reference function it restates (paths relative to the reference repo).
 * Compiled with plain gcc -O2 (no -ffast-math, no FMA contraction), so the
 * hashed Gaussian uses the host glibc log/sqrt exactly as the reference
 * does.
 */
#include "lgh_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) < (b) ? (a) : (b))

/* Host threads for the column-parallel sweeps (orc_set_threads; default 1).
 * Every column of a block solve and of a sparse product is computed by the
 * same operations in the same order whichever thread runs it, so the
 * thread count never changes a result. */
static int g_threads = 1;

void
orc_set_threads (int nt)
{
  g_threads = nt > 1 ? nt : 1;
}

int
orc_get_threads (void)
{
  return g_threads;
}

static double
now_s (void)
{
  struct timespec ts;
  clock_gettime (CLOCK_MONOTONIC, &ts);
  return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

static void *
xcalloc (size_t n, size_t sz)
{
  void *p = calloc (n ? n : 1, sz);
  if (!p) abort ();
  return p;
}

/* ===================================================================== */
/* hashed Gaussian: glr_common_impl.h:35-73 (lgh_randn, lgh_randn_at)     */

double
orc_randn_at (uint64_t seed, int64_t i, int64_t j)
{
  uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (uint64_t) (i + 1)
               + 0xC2B2AE3D27D4EB4FULL * (uint64_t) (j + 1);
  uint64_t s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  s = z ? z : 0x853C49E6748FEA9BULL;
  for (;;) {
    double u1, u2, q;
    uint64_t x;
    x = s; x ^= x >> 12; x ^= x << 25; x ^= x >> 27; s = x;
    u1 = (double) ((x * 2685821657736338717ULL) >> 11) / 9007199254740992.0;
    x = s; x ^= x >> 12; x ^= x << 25; x ^= x >> 27; s = x;
    u2 = (double) ((x * 2685821657736338717ULL) >> 11) / 9007199254740992.0;
    u1 = 2. * u1 - 1.; u2 = 2. * u2 - 1.;
    q = u1 * u1 + u2 * u2;
    if (q > 0. && q < 1.) {
      double f = sqrt (-2. * log (q) / q);
      return u1 * f;          /* the spare u2*f is discarded (:49-50, :71) */
    }
  }
}

/* glr_common_impl.h:441-443: qa[i + j*lda] = lgh_randn_at(seed, rstart+i, j) */
void
orc_randn_fill (uint64_t seed, int64_t row0, int64_t col0, int64_t nr,
                int64_t nc, double *out, int64_t ld)
{
  for (int64_t c = 0; c < nc; c++)
    for (int64_t r = 0; r < nr; r++)
      out[r + c * ld] = orc_randn_at (seed, row0 + r, col0 + c);
}

/* ===================================================================== */
/* CSR                                                                     */

void
orc_csr_mm (int n, const int *rp, const int *ci, const double *v, int ncols,
            const double *X, int ldx, double *Y, int ldy)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads) if (g_threads > 1 && ncols > 1)
  for (int c = 0; c < ncols; c++) {
    const double *x = X + (size_t) c * ldx;
    double       *y = Y + (size_t) c * ldy;
    for (int i = 0; i < n; i++) {
      double s = 0.;
      for (int k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
      y[i] = s;
    }
  }
}

/* Y = P^T R (P: nf x nc CSR) — MatTransposeMatMult at zsolve_impl.h:469 */
static void
csr_tmm (int nf, int nc, const int *rp, const int *ci, const double *v,
         int ncols, const double *R, int ldr, double *Y, int ldy)
{
  for (int c = 0; c < ncols; c++) {
    double *y = Y + (size_t) c * ldy;
    const double *r = R + (size_t) c * ldr;
    for (int j = 0; j < nc; j++) y[j] = 0.;
    for (int i = 0; i < nf; i++)
      for (int k = rp[i]; k < rp[i + 1]; k++) y[ci[k]] += v[k] * r[i];
  }
}

static void
csr_diag_inv (int n, const int *rp, const int *ci, const double *v,
              double *dinv)
{
  for (int i = 0; i < n; i++) {
    double d = 0.;
    for (int k = rp[i]; k < rp[i + 1]; k++) if (ci[k] == i) d += v[k];
    dinv[i] = 1. / d;
  }
}

/* ===================================================================== */
/* dense kernels (the LAPACK/BLAS entry points of f77_decls.h:20-24)      */

void
orc_gemm_tn (int m, int n, int k, const double *A, int lda, const double *B,
             int ldb, double *C, int ldc)
{
  /* C = A^T B, A: k x m, B: k x n */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      const double *a = A + (size_t) i * lda, *b = B + (size_t) j * ldb;
      double s = 0.;
      for (int r = 0; r < k; r++) s += a[r] * b[r];
      C[i + (size_t) j * ldc] = s;
    }
}

void
orc_gemm_nn (int m, int n, int k, double alpha, const double *A, int lda,
             const double *B, int ldb, double beta, double *C, int ldc)
{
  /* C = alpha A B + beta C, A: m x k, B: k x n */
  for (int j = 0; j < n; j++) {
    double *c = C + (size_t) j * ldc;
    if (beta == 0.) for (int i = 0; i < m; i++) c[i] = 0.;
    else if (beta != 1.) for (int i = 0; i < m; i++) c[i] *= beta;
    for (int r = 0; r < k; r++) {
      const double br = alpha * B[r + (size_t) j * ldb];
      const double *a = A + (size_t) r * lda;
      if (br == 0.) continue;
      for (int i = 0; i < m; i++) c[i] += a[i] * br;
    }
  }
}

int
orc_potrf_l (int n, double *G, int ld)
{
  for (int j = 0; j < n; j++) {
    double d = G[j + (size_t) j * ld];
    for (int k = 0; k < j; k++) d -= G[j + (size_t) k * ld] * G[j + (size_t) k * ld];
    if (!(d > 0.)) return j + 1;
    d = sqrt (d);
    G[j + (size_t) j * ld] = d;
    for (int i = j + 1; i < n; i++) {
      double s = G[i + (size_t) j * ld];
      for (int k = 0; k < j; k++) s -= G[i + (size_t) k * ld] * G[j + (size_t) k * ld];
      G[i + (size_t) j * ld] = s / d;
    }
  }
  return 0;
}

int
orc_potrf_u (int n, double *G, int ld)
{
  for (int j = 0; j < n; j++) {
    double d = G[j + (size_t) j * ld];
    for (int k = 0; k < j; k++) d -= G[k + (size_t) j * ld] * G[k + (size_t) j * ld];
    if (!(d > 0.)) return j + 1;
    d = sqrt (d);
    G[j + (size_t) j * ld] = d;
    for (int i = j + 1; i < n; i++) {
      double s = G[j + (size_t) i * ld];
      for (int k = 0; k < j; k++) s -= G[k + (size_t) j * ld] * G[k + (size_t) i * ld];
      G[j + (size_t) i * ld] = s / d;
    }
  }
  return 0;
}

void
orc_trsm_rltn (int m, int n, const double *L, int ldl, double *X, int ldx)
{
  /* solve Y L^T = X column by column: Y_j = (X_j - sum_k<j L_jk Y_k)/L_jj */
  for (int j = 0; j < n; j++) {
    double *xj = X + (size_t) j * ldx;
    for (int k = 0; k < j; k++) {
      const double l = L[j + (size_t) k * ldl];
      const double *yk = X + (size_t) k * ldx;
      if (l != 0.) for (int i = 0; i < m; i++) xj[i] -= l * yk[i];
    }
    {
      const double inv = 1. / L[j + (size_t) j * ldl];
      for (int i = 0; i < m; i++) xj[i] *= inv;
    }
  }
}

void
orc_trsm_runn (int m, int n, const double *R, int ldr, double *X, int ldx)
{
  /* solve Y R = X: Y_j = (X_j - sum_k<j Y_k R_kj)/R_jj */
  for (int j = 0; j < n; j++) {
    double *xj = X + (size_t) j * ldx;
    for (int k = 0; k < j; k++) {
      const double r = R[k + (size_t) j * ldr];
      const double *yk = X + (size_t) k * ldx;
      if (r != 0.) for (int i = 0; i < m; i++) xj[i] -= r * yk[i];
    }
    {
      const double inv = 1. / R[j + (size_t) j * ldr];
      for (int i = 0; i < m; i++) xj[i] *= inv;
    }
  }
}

/* Symmetric eigensolver: Householder tridiagonalization with explicit Q,
 * then implicit symmetric QR steps with the Wilkinson shift (Golub & Van
 * Loan, Matrix Computations, Alg. 8.3.1 and 8.3.3), eigenvectors
 * accumulated by the Givens rotations; ascending order at the end.      */
static int
tridiag_qr (int n, double *d, double *e, double *Z, int ldz)
{
  const double eps = DBL_EPSILON;
  int m = n - 1, iter = 0;
  while (m > 0) {
    int l;
    if (fabs (e[m - 1]) <= eps * (fabs (d[m - 1]) + fabs (d[m]))) {
      e[m - 1] = 0.; m--; iter = 0; continue;
    }
    if (++iter > 60) return m;
    l = m - 1;
    while (l > 0 && fabs (e[l - 1]) > eps * (fabs (d[l - 1]) + fabs (d[l])))
      l--;
    if (l > 0) e[l - 1] = 0.;
    {
      const double dd = 0.5 * (d[m - 1] - d[m]);
      const double em = e[m - 1];
      const double den = dd + (dd >= 0. ? hypot (dd, em) : -hypot (dd, em));
      const double mu = d[m] - (den != 0. ? em * em / den : 0.);
      double x = d[l] - mu, z = e[l];
      for (int k = l; k < m; k++) {
        double r = hypot (x, z), c = 1., s = 0.;
        if (r != 0.) { c = x / r; s = z / r; }
        if (k > l) e[k - 1] = r;
        {
          const double a = d[k], b = e[k], cc = d[k + 1];
          d[k] = c * c * a + 2. * c * s * b + s * s * cc;
          d[k + 1] = s * s * a - 2. * c * s * b + c * c * cc;
          e[k] = c * s * (cc - a) + (c * c - s * s) * b;
        }
        if (k + 1 < m) {
          const double f = e[k + 1];
          z = s * f;
          e[k + 1] = c * f;
        }
        x = e[k];
        for (int i = 0; i < n; i++) {
          const double zk = Z[i + (size_t) k * ldz];
          const double zk1 = Z[i + (size_t) (k + 1) * ldz];
          Z[i + (size_t) k * ldz] = c * zk + s * zk1;
          Z[i + (size_t) (k + 1) * ldz] = -s * zk + c * zk1;
        }
      }
    }
  }
  return 0;
}

int
orc_syev (int n, double *A, int lda, double *w)
{
  double *Q = xcalloc ((size_t) n * n, sizeof (double));
  double *V = xcalloc ((size_t) n * n, sizeof (double));
  double *betas = xcalloc ((size_t) n, sizeof (double));
  double *d = xcalloc ((size_t) n, sizeof (double));
  double *e = xcalloc ((size_t) n, sizeof (double));
  double *p = xcalloc ((size_t) n, sizeof (double));
  double *wv = xcalloc ((size_t) n, sizeof (double));
  double *T = xcalloc ((size_t) n * n, sizeof (double));
  int info;

  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) T[i + (size_t) j * n] = A[i + (size_t) j * lda];
  /* Householder: zero T[k+2:, k] for k = 0..n-3; v_k stored in V[:, k] */
  for (int k = 0; k + 2 < n; k++) {
    const int m = n - k - 1;
    double *x = T + (k + 1) + (size_t) k * n;
    double *v = V + (k + 1) + (size_t) k * n;
    double nrm = 0., alpha, vtv = 0., K = 0.;
    for (int i = 0; i < m; i++) nrm += x[i] * x[i];
    nrm = sqrt (nrm);
    if (nrm == 0.) { betas[k] = 0.; continue; }
    alpha = (x[0] > 0.) ? -nrm : nrm;
    for (int i = 0; i < m; i++) v[i] = x[i];
    v[0] -= alpha;
    for (int i = 0; i < m; i++) vtv += v[i] * v[i];
    if (vtv == 0.) { betas[k] = 0.; continue; }
    betas[k] = 2. / vtv;
    /* p = beta * T22 v, T22 = T[k+1:, k+1:] */
    for (int i = 0; i < m; i++) {
      double s = 0.;
      for (int j = 0; j < m; j++)
        s += T[(k + 1 + i) + (size_t) (k + 1 + j) * n] * v[j];
      p[i] = betas[k] * s;
    }
    for (int i = 0; i < m; i++) K += p[i] * v[i];
    K *= 0.5 * betas[k];
    for (int i = 0; i < m; i++) wv[i] = p[i] - K * v[i];
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++)
        T[(k + 1 + i) + (size_t) (k + 1 + j) * n] -= v[i] * wv[j] + wv[i] * v[j];
    T[(k + 1) + (size_t) k * n] = alpha;
    T[k + (size_t) (k + 1) * n] = alpha;
    for (int i = 1; i < m; i++) {
      T[(k + 1 + i) + (size_t) k * n] = 0.;
      T[k + (size_t) (k + 1 + i) * n] = 0.;
    }
  }
  for (int i = 0; i < n; i++) {
    d[i] = T[i + (size_t) i * n];
    e[i] = (i + 1 < n) ? T[(i + 1) + (size_t) i * n] : 0.;
  }
  /* Q = H_0 H_1 ... H_{n-3}, accumulated backwards */
  for (int i = 0; i < n; i++) Q[i + (size_t) i * n] = 1.;
  for (int k = n - 3; k >= 0; k--) {
    const int m = n - k - 1;
    const double *v = V + (k + 1) + (size_t) k * n;
    if (betas[k] == 0.) continue;
    for (int j = k + 1; j < n; j++) {
      double *q = Q + (k + 1) + (size_t) j * n;
      double s = 0.;
      for (int i = 0; i < m; i++) s += v[i] * q[i];
      s *= betas[k];
      for (int i = 0; i < m; i++) q[i] -= s * v[i];
    }
  }
  info = tridiag_qr (n, d, e, Q, n);
  /* ascending selection sort, permuting vectors */
  for (int i = 0; i < n; i++) {
    int kmin = i;
    for (int j = i + 1; j < n; j++) if (d[j] < d[kmin]) kmin = j;
    if (kmin != i) {
      double t = d[i]; d[i] = d[kmin]; d[kmin] = t;
      for (int r = 0; r < n; r++) {
        double q = Q[r + (size_t) i * n];
        Q[r + (size_t) i * n] = Q[r + (size_t) kmin * n];
        Q[r + (size_t) kmin * n] = q;
      }
    }
  }
  for (int i = 0; i < n; i++) w[i] = d[i];
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) A[i + (size_t) j * lda] = Q[i + (size_t) j * n];
  free (Q); free (V); free (betas); free (d); free (e); free (p); free (wv);
  free (T);
  return info;
}

/* eigenvalues of a symmetric tridiagonal (Lanczos matrix of CG) */
static void
tridiag_eigvals (int n, const double *diag, const double *off, double *emin,
                 double *emax)
{
  double *d = xcalloc ((size_t) n, sizeof (double));
  double *e = xcalloc ((size_t) n, sizeof (double));
  double *Z = xcalloc ((size_t) n * n, sizeof (double));
  for (int i = 0; i < n; i++) { d[i] = diag[i]; e[i] = (i + 1 < n) ? off[i] : 0.; }
  for (int i = 0; i < n; i++) Z[i + (size_t) i * n] = 1.;
  tridiag_qr (n, d, e, Z, n);
  *emin = d[0]; *emax = d[0];
  for (int i = 1; i < n; i++) { *emin = MIN (*emin, d[i]); *emax = MAX (*emax, d[i]); }
  free (d); free (e); free (Z);
}

/* dense LU with partial pivoting (MatLUFactor at zsolve_impl.h:234-236) */
static int
lu_factor (int n, double *A, int *piv)
{
  for (int k = 0; k < n; k++) {
    int p = k;
    double amax = fabs (A[k + (size_t) k * n]);
    for (int i = k + 1; i < n; i++)
      if (fabs (A[i + (size_t) k * n]) > amax) { amax = fabs (A[i + (size_t) k * n]); p = i; }
    piv[k] = p;
    if (amax == 0.) return k + 1;
    if (p != k)
      for (int j = 0; j < n; j++) {
        double t = A[k + (size_t) j * n];
        A[k + (size_t) j * n] = A[p + (size_t) j * n];
        A[p + (size_t) j * n] = t;
      }
    for (int i = k + 1; i < n; i++) A[i + (size_t) k * n] /= A[k + (size_t) k * n];
    for (int j = k + 1; j < n; j++) {
      const double akj = A[k + (size_t) j * n];
      if (akj == 0.) continue;
      for (int i = k + 1; i < n; i++) A[i + (size_t) j * n] -= A[i + (size_t) k * n] * akj;
    }
  }
  return 0;
}

static void
lu_solve (int n, const double *LU, const int *piv, double *b)
{
  for (int k = 0; k < n; k++) {
    if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
  }
  for (int k = 0; k < n; k++)
    for (int i = k + 1; i < n; i++) b[i] -= LU[i + (size_t) k * n] * b[k];
  for (int k = n - 1; k >= 0; k--) {
    b[k] /= LU[k + (size_t) k * n];
    for (int i = 0; i < k; i++) b[i] -= LU[i + (size_t) k * n] * b[k];
  }
}

/* ===================================================================== */
/* Chebyshev core: zsolve_impl.h:104-141.  Contract: on entry R = B - A X. */

typedef void (*pc_fn) (void *ctx, int ncols, const double *R, int ldr,
                       double *Z, int ldz);

typedef struct {
  int n; const int *rp, *ci; const double *v;
} csr_t;

static void
cheb_core (const csr_t *A, pc_fn pcap, void *pcctx, double lmin, double lmax,
           int nit, int ncols, double *X, double *R, double *Z, double *D,
           double *AD, int ld)
{
  const double theta = 0.5 * (lmax + lmin);
  const double delta = 0.5 * (lmax - lmin);
  const double sigma1 = theta / delta;
  double rho_old = 1.0 / sigma1, rho;
  const size_t tot = (size_t) ld * ncols;
  const int n = A->n;

  pcap (pcctx, ncols, R, ld, Z, ld);
  for (size_t q = 0; q < tot; q++) D[q] = Z[q];
  {
    const double s = 1.0 / theta;
    for (int c = 0; c < ncols; c++)
      for (int i = 0; i < n; i++) D[i + (size_t) c * ld] *= s;
  }
  for (int c = 0; c < ncols; c++)
    for (int i = 0; i < n; i++) X[i + (size_t) c * ld] += D[i + (size_t) c * ld];
  orc_csr_mm (n, A->rp, A->ci, A->v, ncols, D, ld, AD, ld);
  for (int c = 0; c < ncols; c++)
    for (int i = 0; i < n; i++) R[i + (size_t) c * ld] -= AD[i + (size_t) c * ld];
  for (int k = 2; k <= nit; k++) {
    pcap (pcctx, ncols, R, ld, Z, ld);
    rho = 1.0 / (2.0 * sigma1 - rho_old);
    {
      const double a = rho * rho_old, b = 2.0 * rho / delta;
      for (int c = 0; c < ncols; c++)
        for (int i = 0; i < n; i++) {
          const size_t q = i + (size_t) c * ld;
          D[q] *= a;
          D[q] += b * Z[q];
          X[q] += D[q];
        }
    }
    orc_csr_mm (n, A->rp, A->ci, A->v, ncols, D, ld, AD, ld);
    for (int c = 0; c < ncols; c++)
      for (int i = 0; i < n; i++) R[i + (size_t) c * ld] -= AD[i + (size_t) c * ld];
    rho_old = rho;
  }
}

/* row-scaling PC: zsolve_impl.h:143-152 */
typedef struct { int n; const double *dinv; } rowscale_ctx;

static void
rowscale_pc (void *vctx, int ncols, const double *R, int ldr, double *Z,
             int ldz)
{
  rowscale_ctx *c = (rowscale_ctx *) vctx;
  for (int j = 0; j < ncols; j++)
    for (int i = 0; i < c->n; i++)
      Z[i + (size_t) j * ldz] = R[i + (size_t) j * ldr] * c->dinv[i];
}

/* zsolve_impl.h:93-102 */
int
orc_cheb_count (double lmin, double lmax, double tol)
{
  const double kap = lmax / lmin;
  const double rho = (sqrt (kap) - 1.) / (sqrt (kap) + 1.);
  return (int) ceil (log (2. / tol) / log (1. / rho));
}

void
orc_cheb_jacobi (int n, const int *rp, const int *ci, const double *v,
                 const double *dinv, double lmin, double lmax, int nit,
                 int ncols, const double *B, int ldb, double *X, int ldx)
{
  /* blocked asolve contract (zsolve_impl.h:629-633): X = 0, R = B */
  csr_t A = { n, rp, ci, v };
  rowscale_ctx pc = { n, dinv };
  const size_t tot = (size_t) n * ncols;
  double *Xw = xcalloc (tot, sizeof (double));
  double *R = xcalloc (tot, sizeof (double));
  double *Z = xcalloc (tot, sizeof (double));
  double *D = xcalloc (tot, sizeof (double));
  double *AD = xcalloc (tot, sizeof (double));
  for (int c = 0; c < ncols; c++)
    memcpy (R + (size_t) c * n, B + (size_t) c * ldb, sizeof (double) * n);
  cheb_core (&A, rowscale_pc, &pc, lmin, lmax, nit, ncols, Xw, R, Z, D, AD, n);
  for (int c = 0; c < ncols; c++)
    memcpy (X + (size_t) c * ldx, Xw + (size_t) c * n, sizeof (double) * n);
  free (Xw); free (R); free (Z); free (D); free (AD);
}

/* ===================================================================== */
/* multigrid hierarchy + block V-cycle: zsolve_impl.h:38-65, 385-495     */

typedef struct {
  int n, nc;                 /* this level's size; level l-1's size       */
  int *rp, *ci; double *v;   /* A_l                                       */
  int *prp, *pci; double *pv;/* P_l: n x nc (levels > 0)                  */
  double *dinv;
  double slo, shi; int nu;
  /* dense LU of A_0 */
  double *lu; int *piv;
} mglev_t;

struct orc_mg {
  int nl, tile;
  int gamma;                    /* cycle index: 1 V-cycle, 2 W-cycle */
  mglev_t *lev;
};

orc_mg_t *
orc_mg_create (int nl, int tile)
{
  orc_mg_t *h = xcalloc (1, sizeof (*h));
  h->nl = nl; h->tile = tile; h->gamma = 1;
  h->lev = xcalloc ((size_t) nl, sizeof (mglev_t));
  return h;
}

static int *
dup_i (const int *a, size_t n)
{
  int *b = xcalloc (n, sizeof (int));
  memcpy (b, a, n * sizeof (int));
  return b;
}

static double *
dup_d (const double *a, size_t n)
{
  double *b = xcalloc (n, sizeof (double));
  memcpy (b, a, n * sizeof (double));
  return b;
}

int
orc_mg_set_level (orc_mg_t *h, int l, int n, const int *rp, const int *ci,
                  const double *v, int nc_prev, const int *prp,
                  const int *pci, const double *pv, double slo, double shi,
                  int nu)
{
  mglev_t *lv = &h->lev[l];
  lv->n = n;
  lv->rp = dup_i (rp, (size_t) n + 1);
  lv->ci = dup_i (ci, (size_t) rp[n]);
  lv->v = dup_d (v, (size_t) rp[n]);
  lv->dinv = xcalloc ((size_t) n, sizeof (double));
  csr_diag_inv (n, lv->rp, lv->ci, lv->v, lv->dinv);
  lv->slo = slo; lv->shi = shi; lv->nu = nu;
  if (l > 0) {
    lv->nc = nc_prev;
    lv->prp = dup_i (prp, (size_t) n + 1);
    lv->pci = dup_i (pci, (size_t) prp[n]);
    lv->pv = dup_d (pv, (size_t) prp[n]);
  }
  else {
    /* dense LU once (zsolve_impl.h:232-237) */
    lv->lu = xcalloc ((size_t) n * n, sizeof (double));
    lv->piv = xcalloc ((size_t) n, sizeof (int));
    for (int i = 0; i < n; i++)
      for (int k = rp[i]; k < rp[i + 1]; k++)
        lv->lu[i + (size_t) ci[k] * n] += v[k];
    return lu_factor (n, lv->lu, lv->piv);
  }
  return 0;
}

void
orc_mg_free (orc_mg_t *h)
{
  if (!h) return;
  for (int l = 0; l < h->nl; l++) {
    mglev_t *lv = &h->lev[l];
    free (lv->rp); free (lv->ci); free (lv->v); free (lv->dinv);
    free (lv->prp); free (lv->pci); free (lv->pv); free (lv->lu); free (lv->piv);
  }
  free (h->lev);
  free (h);
}

/* correction-form smoothing (zsolve_impl.h:385-404):
 * D = X; AD = A D; R = B - AD; cheb_core(A, rowscale, slo, shi, nu) */
static void
mg_smooth (mglev_t *lv, int ncols, const double *B, double *X, double *R,
           double *Z, double *D, double *AD)
{
  const int n = lv->n;
  csr_t A = { n, lv->rp, lv->ci, lv->v };
  rowscale_ctx pc = { n, lv->dinv };
  memcpy (D, X, sizeof (double) * (size_t) n * ncols);
  orc_csr_mm (n, lv->rp, lv->ci, lv->v, ncols, D, n, AD, n);
  for (size_t q = 0; q < (size_t) n * ncols; q++) R[q] = B[q] - AD[q];
  cheb_core (&A, rowscale_pc, &pc, lv->slo, lv->shi, lv->nu, ncols, X, R, Z,
             D, AD, n);
}

void
orc_mg_set_cycle (orc_mg_t *h, int gamma)
{
  h->gamma = gamma;
}

/* zsolve_impl.h:406-485; B, X are n_l x ncols, ld n_l.  X holds the
 * initial guess (zero for a plain cycle): the correction-form
 * pre-smoother starts from B - A X.  With gamma = 2 a level above the
 * coarsest corrects twice, the second cycle starting from the first
 * one's result (W-cycle; the cycle index is the engine's choice, the
 * reference's PCMG cycle is the V-cycle) */
static void
mg_run (orc_mg_t *h, int l, int ncols, const double *B, double *X)
{
  mglev_t *lv = &h->lev[l];
  const int n = lv->n;
  if (l == 0) {
    for (int c = 0; c < ncols; c++) {
      double *x = X + (size_t) c * n;
      memcpy (x, B + (size_t) c * n, sizeof (double) * n);
      lu_solve (n, lv->lu, lv->piv, x);
    }
    return;
  }
  {
    const size_t tot = (size_t) n * ncols;
    const int nc = lv->nc;
    double *R = xcalloc (tot, sizeof (double));
    double *Z = xcalloc (tot, sizeof (double));
    double *D = xcalloc (tot, sizeof (double));
    double *AD = xcalloc (tot, sizeof (double));
    double *Bc = xcalloc ((size_t) nc * ncols, sizeof (double));
    double *Xc = xcalloc ((size_t) nc * ncols, sizeof (double));
    mg_smooth (lv, ncols, B, X, R, Z, D, AD);                     /* pre  */
    memcpy (D, X, sizeof (double) * tot);
    orc_csr_mm (n, lv->rp, lv->ci, lv->v, ncols, D, n, AD, n);
    for (size_t q = 0; q < tot; q++) R[q] = B[q] - AD[q];
    csr_tmm (n, nc, lv->prp, lv->pci, lv->pv, ncols, R, n, Bc, nc);
    mg_run (h, l - 1, ncols, Bc, Xc);
    if (h->gamma == 2 && l - 1 >= 1) mg_run (h, l - 1, ncols, Bc, Xc);
    orc_csr_mm (n, lv->prp, lv->pci, lv->pv, ncols, Xc, nc, AD, n); /* P X_c */
    for (size_t q = 0; q < tot; q++) X[q] += AD[q];
    mg_smooth (lv, ncols, B, X, R, Z, D, AD);                     /* post */
    free (R); free (Z); free (D); free (AD); free (Bc); free (Xc);
  }
}

void
orc_mg_vcycle (orc_mg_t *h, int ncols, const double *B, int ldb, double *X,
               int ldx)
{
  const int n = h->lev[h->nl - 1].n;
  double *Bw = xcalloc ((size_t) n * ncols, sizeof (double));
  double *Xw = xcalloc ((size_t) n * ncols, sizeof (double));
  for (int c = 0; c < ncols; c++)
    memcpy (Bw + (size_t) c * n, B + (size_t) c * ldb, sizeof (double) * n);
  mg_run (h, h->nl - 1, ncols, Bw, Xw);
  for (int c = 0; c < ncols; c++)
    memcpy (X + (size_t) c * ldx, Xw + (size_t) c * n, sizeof (double) * n);
  free (Bw); free (Xw);
}

static void
vcycle_pc (void *vctx, int ncols, const double *R, int ldr, double *Z,
           int ldz)
{
  orc_mg_vcycle ((orc_mg_t *) vctx, ncols, R, ldr, Z, ldz);
}

void
orc_cheb_mg (orc_mg_t *h, double lmin, double lmax, int nit, int ncols,
             const double *B, int ldb, double *X, int ldx)
{
  mglev_t *top = &h->lev[h->nl - 1];
  const int n = top->n;
  csr_t A = { n, top->rp, top->ci, top->v };
  const size_t tot = (size_t) n * ncols;
  double *Xw = xcalloc (tot, sizeof (double));
  double *R = xcalloc (tot, sizeof (double));
  double *Z = xcalloc (tot, sizeof (double));
  double *D = xcalloc (tot, sizeof (double));
  double *AD = xcalloc (tot, sizeof (double));
  for (int c = 0; c < ncols; c++)
    memcpy (R + (size_t) c * n, B + (size_t) c * ldb, sizeof (double) * n);
  cheb_core (&A, vcycle_pc, h, lmin, lmax, nit, ncols, Xw, R, Z, D, AD, n);
  for (int c = 0; c < ncols; c++)
    memcpy (X + (size_t) c * ldx, Xw + (size_t) c * n, sizeof (double) * n);
  free (Xw); free (R); free (Z); free (D); free (AD);
}

/* ===================================================================== */
/* PCG with Lanczos extreme eigenvalue estimates (KSPCG +
 * KSPComputeExtremeSingularValues, used at zsolve_impl.h:178-196,
 * 284-313, 569-597; convergence on the preconditioned residual norm).   */

int
orc_cg (int n, const int *rp, const int *ci, const double *v, int pc,
        orc_mg_t *h, const double *b, double *x, double rtol, int maxit,
        double *emin, double *emax)
{
  double *r = xcalloc ((size_t) n, sizeof (double));
  double *z = xcalloc ((size_t) n, sizeof (double));
  double *p = xcalloc ((size_t) n, sizeof (double));
  double *q = xcalloc ((size_t) n, sizeof (double));
  double *dinv = xcalloc ((size_t) n, sizeof (double));
  double *alph = xcalloc ((size_t) maxit + 1, sizeof (double));
  double *bet = xcalloc ((size_t) maxit + 1, sizeof (double));
  double dp, dp_new, z0 = 0., zn;
  int it = 0;

  csr_diag_inv (n, rp, ci, v, dinv);
  for (int i = 0; i < n; i++) { x[i] = 0.; r[i] = b[i]; }
  if (pc == 1) orc_mg_vcycle (h, 1, r, n, z, n);
  else for (int i = 0; i < n; i++) z[i] = r[i] * dinv[i];
  dp = 0.; for (int i = 0; i < n; i++) { dp += r[i] * z[i]; z0 += z[i] * z[i]; }
  z0 = sqrt (z0);
  for (int i = 0; i < n; i++) p[i] = z[i];
  if (z0 > 0.) {
    for (it = 0; it < maxit;) {
      double pq = 0., a, beta;
      orc_csr_mm (n, rp, ci, v, 1, p, n, q, n);
      for (int i = 0; i < n; i++) pq += p[i] * q[i];
      a = dp / pq;
      for (int i = 0; i < n; i++) { x[i] += a * p[i]; r[i] -= a * q[i]; }
      if (pc == 1) orc_mg_vcycle (h, 1, r, n, z, n);
      else for (int i = 0; i < n; i++) z[i] = r[i] * dinv[i];
      dp_new = 0.; zn = 0.;
      for (int i = 0; i < n; i++) { dp_new += r[i] * z[i]; zn += z[i] * z[i]; }
      beta = dp_new / dp;
      alph[it] = a; bet[it] = beta;
      it++;
      if (sqrt (zn) <= rtol * z0) break;
      for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
      dp = dp_new;
    }
  }
  if (emin && emax) {
    if (it == 0) { *emin = *emax = 1.; }
    else {
      double *dg = xcalloc ((size_t) it, sizeof (double));
      double *of = xcalloc ((size_t) it, sizeof (double));
      for (int k = 0; k < it; k++) {
        dg[k] = 1. / alph[k] + (k > 0 ? bet[k - 1] / alph[k - 1] : 0.);
        of[k] = sqrt (bet[k]) / alph[k];
      }
      tridiag_eigvals (it, dg, of, emin, emax);
      free (dg); free (of);
    }
  }
  free (r); free (z); free (p); free (q); free (dinv); free (alph); free (bet);
  return it;
}

/* ===================================================================== */
/* Z-solvers (the prior's callbacks: glr_common_impl.h:170-234,
 * zsolve_impl.h:664-715)                                                 */

struct orc_zs {
  int kind;                  /* 0 diag, 1 mat                            */
  int n;
  double *zdiag;
  const int *rp, *ci; const double *v;
  double *dinv;
  orc_mg_t *h;
  int mode, nit;
  double lmin, lmax, ksp_rtol;
};

orc_zs_t *
orc_zs_diag (int n, const double *zdiag)
{
  orc_zs_t *z = xcalloc (1, sizeof (*z));
  z->kind = 0; z->n = n;
  z->zdiag = dup_d (zdiag, (size_t) n);
  return z;
}

orc_zs_t *
orc_zs_mat (int n, const int *rp, const int *ci, const double *v,
            orc_mg_t *h, int mode, double lmin, double lmax, int nit,
            double ksp_rtol)
{
  orc_zs_t *z = xcalloc (1, sizeof (*z));
  z->kind = 1; z->n = n; z->rp = rp; z->ci = ci; z->v = v; z->h = h;
  z->mode = mode; z->lmin = lmin; z->lmax = lmax; z->nit = nit;
  z->ksp_rtol = ksp_rtol;
  z->dinv = xcalloc ((size_t) n, sizeof (double));
  csr_diag_inv (n, rp, ci, v, z->dinv);
  return z;
}

void
orc_zs_free (orc_zs_t *zs)
{
  if (!zs) return;
  free (zs->zdiag); free (zs->dinv); free (zs);
}

void
orc_zs_apply (orc_zs_t *zs, int ncols, const double *X, int ldx, double *Y,
              int ldy)
{
  if (zs->kind == 0) {
    for (int c = 0; c < ncols; c++)
      for (int i = 0; i < zs->n; i++)
        Y[i + (size_t) c * ldy] = X[i + (size_t) c * ldx] * zs->zdiag[i];
    return;
  }
  orc_csr_mm (zs->n, zs->rp, zs->ci, zs->v, ncols, X, ldx, Y, ldy);
}

void
orc_zs_solve_vec (orc_zs_t *zs, const double *b, double *x)
{
  if (zs->kind == 0) {
    for (int i = 0; i < zs->n; i++) x[i] = b[i] / zs->zdiag[i];
    return;
  }
  /* lgh_prior_mat_solveZ: x = 0; KSPSolve (CG + AMG) to ksp_rtol        */
  orc_cg (zs->n, zs->rp, zs->ci, zs->v, zs->h ? 1 : 0, zs->h, b, x,
          zs->ksp_rtol, 10000, NULL, NULL);
}

void
orc_zs_solve_block (orc_zs_t *zs, int ncols, const double *X, int ldx,
                    double *Y, int ldy)
{
  if (zs->kind == 0 || zs->mode < 2) {
    for (int c = 0; c < ncols; c++)
      orc_zs_solve_vec (zs, X + (size_t) c * ldx, Y + (size_t) c * ldy);
    return;
  }
  if (g_threads > 1 && ncols > 1) {
    /* one column per task: the same per-column arithmetic as the block */
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int c = 0; c < ncols; c++) {
      if (zs->mode == 4)
        orc_cheb_jacobi (zs->n, zs->rp, zs->ci, zs->v, zs->dinv, zs->lmin,
                         zs->lmax, zs->nit, 1, X + (size_t) c * ldx, ldx,
                         Y + (size_t) c * ldy, ldy);
      else
        orc_cheb_mg (zs->h, zs->lmin, zs->lmax, zs->nit, 1, X + (size_t) c * ldx,
                     ldx, Y + (size_t) c * ldy, ldy);
    }
    return;
  }
  if (zs->mode == 4)
    orc_cheb_jacobi (zs->n, zs->rp, zs->ci, zs->v, zs->dinv, zs->lmin,
                     zs->lmax, zs->nit, ncols, X, ldx, Y, ldy);
  else
    orc_cheb_mg (zs->h, zs->lmin, zs->lmax, zs->nit, ncols, X, ldx, Y, ldy);
}

/* ===================================================================== */
/* GLR: glr_common_impl.h:252-278 (object), 314-608 (replicated build)    */

struct orc_glr {
  int n;
  int brp_n; const int *brp, *bci; const double *bv;
  const double *mass;
  double *msqrt, *minvsqrt;
  orc_zs_t *zs;
  orc_glr_opts_t opts;
  double *U, *lam;
  int kept;
  /* kept for incremental extension (glr_scalapack_impl.h:84-91)        */
  double *Q, *Araw;
  int ell;
  double *w1, *w2;
};

/* glr_common_impl.h:316-336: Y = Msqrt Z^{-1} B Z^{-T} Msqrt X */
void
orc_glr_apply_F (orc_glr_t *g, int ncols, const double *X, int ldx,
                 double *Y, int ldy)
{
  const int n = g->n;
  double *W = xcalloc ((size_t) n * ncols, sizeof (double));
  double *T = xcalloc ((size_t) n * ncols, sizeof (double));
  for (int c = 0; c < ncols; c++)
    for (int i = 0; i < n; i++)
      W[i + (size_t) c * n] = X[i + (size_t) c * ldx] * g->msqrt[i];
  orc_zs_solve_block (g->zs, ncols, W, n, T, n);              /* Z^{-T}  */
  orc_csr_mm (n, g->brp, g->bci, g->bv, ncols, T, n, W, n);    /* B       */
  orc_zs_solve_block (g->zs, ncols, W, n, T, n);              /* Z^{-1}  */
  for (int c = 0; c < ncols; c++)
    for (int i = 0; i < n; i++)
      Y[i + (size_t) c * ldy] = T[i + (size_t) c * n] * g->msqrt[i];
  free (W); free (T);
}

/* glr_common_impl.h:340-407 (gram + CholQR2 with the one-time jitter)   */
static int
cholqr2 (int n, int ncols, double *X, double *G)
{
  for (int pass = 0; pass < 2; pass++) {
    int info = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
      orc_gemm_tn (ncols, ncols, n, X, n, X, n, G, ncols);
      if (attempt == 1) {
        double tr = 0.;
        for (int i = 0; i < ncols; i++) tr += G[i + (size_t) i * ncols];
        for (int i = 0; i < ncols; i++) G[i + (size_t) i * ncols] += 1e-12 * tr / ncols;
      }
      info = orc_potrf_l (ncols, G, ncols);
      if (info == 0) break;
    }
    if (info != 0) return -1;
    orc_trsm_rltn (n, ncols, G, ncols, X, n);
  }
  return 0;
}

/* glr_scalapack_impl.h:110-148 */
static int
panel_cholqr (int n, double *pa, int bw)
{
  double *G = xcalloc ((size_t) bw * bw, sizeof (double));
  for (int pass = 0; pass < 2; pass++) {
    int info;
    orc_gemm_tn (bw, bw, n, pa, n, pa, n, G, bw);
    info = orc_potrf_u (bw, G, bw);
    if (info != 0) {
      double dmax = 0.;
      orc_gemm_tn (bw, bw, n, pa, n, pa, n, G, bw);
      for (int i = 0; i < bw; i++) dmax = MAX (dmax, G[i + (size_t) i * bw]);
      for (int i = 0; i < bw; i++) G[i + (size_t) i * bw] += 1.e-12 * MAX (dmax, 1.);
      info = orc_potrf_u (bw, G, bw);
      if (info != 0) { free (G); return -1; }
    }
    for (int j = 0; j < bw; j++)
      for (int i = j + 1; i < bw; i++) G[i + (size_t) j * bw] = 0.;
    orc_trsm_runn (n, bw, G, bw, pa, n);
  }
  free (G);
  return 0;
}

/* glr_scalapack_impl.h:153-195 */
static int
bcgs2 (int n, double *Y, int ell, int b, int p_start)
{
  double *C = xcalloc ((size_t) ell * b, sizeof (double));
  for (int p0 = p_start; p0 < ell; p0 += b) {
    const int bw = MIN (b, ell - p0);
    double *pa = Y + (size_t) p0 * n;
    for (int sweep = 0; sweep < 2; sweep++) {
      if (p0 > 0) {
        orc_gemm_tn (p0, bw, n, Y, n, pa, n, C, p0);
        orc_gemm_nn (n, bw, p0, -1.0, Y, n, C, p0, 1.0, pa, n);
      }
      if (panel_cholqr (n, pa, bw) != 0) { free (C); return -1; }
    }
  }
  free (C);
  return 0;
}

/* selection: glr_common_impl.h:517-536 / glr_scalapack_impl.h:375-425.
 * T holds eigenvectors (ell x ell), evals ascending.  Fills lam, kept,
 * U = Q C_kept, report.                                                  */
static void
select_and_form_U (orc_glr_t *g, int ell, const double *Q, const double *T,
                   const double *evals, orc_glr_report_t *rep, int **order_out)
{
  const int n = g->n;
  int *order = xcalloc ((size_t) ell, sizeof (int));
  double absmax, cut;
  int kept = 0;
  for (int i = 0; i < ell; i++) order[i] = i;
  for (int i = 0; i < ell; i++)
    for (int j = i + 1; j < ell; j++)
      if (fabs (evals[order[j]]) > fabs (evals[order[i]])) {
        int t = order[i]; order[i] = order[j]; order[j] = t;
      }
  absmax = fabs (evals[order[0]]);
  cut = 0.;
  if (g->opts.trunc_abs > 0.) cut = MAX (cut, g->opts.trunc_abs);
  if (g->opts.trunc_rel > 0.) cut = MAX (cut, g->opts.trunc_rel * absmax);
  while (kept < ell && fabs (evals[order[kept]]) > cut) kept++;
  rep->kept = kept;
  rep->lam_abs_max = absmax;
  rep->next_abs = (kept < ell) ? fabs (evals[order[kept]]) : 0.;
  rep->lam_raw_max = evals[ell - 1];
  rep->lam_raw_min = evals[0];
  rep->n_negative_raw = 0;
  for (int i = 0; i < ell; i++) if (evals[i] < 0.) rep->n_negative_raw++;

  free (g->lam); free (g->U);
  g->lam = xcalloc ((size_t) MAX (kept, 1), sizeof (double));
  g->U = xcalloc ((size_t) n * MAX (kept, 1), sizeof (double));
  {
    double *C = xcalloc ((size_t) ell * MAX (kept, 1), sizeof (double));
    for (int j = 0; j < kept; j++) {
      const double l = evals[order[j]];
      for (int i = 0; i < ell; i++) C[i + (size_t) j * ell] = T[i + (size_t) order[j] * ell];
      switch (g->opts.treat) {
      case 0: g->lam[j] = fabs (l); break;
      case 1: g->lam[j] = MAX (l, 0.); break;
      default: g->lam[j] = l; break;
      }
    }
    if (kept > 0) orc_gemm_nn (n, kept, ell, 1.0, Q, n, C, ell, 0.0, g->U, n);
    free (C);
  }
  g->kept = kept;
  *order_out = order;
}

/* glr_common_impl.h:576-600 */
static void
residual_check (orc_glr_t *g, const double *evals, const int *order,
                orc_glr_report_t *rep, double *t_op)
{
  const int n = g->n, k = g->kept;
  double *FU, t0;
  if (!g->opts.check || k == 0) return;
  FU = xcalloc ((size_t) n * k, sizeof (double));
  t0 = now_s ();
  orc_glr_apply_F (g, k, g->U, n, FU, n);
  *t_op += now_s () - t0;
  for (int j = 0; j < k; j++) {
    const double lraw = evals[order[j]];
    double s = 0.;
    for (int i = 0; i < n; i++) {
      const double r = FU[i + (size_t) j * n] - lraw * g->U[i + (size_t) j * n];
      s += r * r;
    }
    rep->resid_max = MAX (rep->resid_max, sqrt (s));
  }
  free (FU);
}

/* symmetrize (glr_common_impl.h:490-494) + eig; keeps raw A for extend  */
static int
eig_of_raw (int ell, const double *Araw, double *T, double *evals)
{
  for (int j = 0; j < ell; j++)
    for (int i = 0; i < ell; i++) T[i + (size_t) j * ell] = Araw[i + (size_t) j * ell];
  for (int j = 0; j < ell; j++)
    for (int i = 0; i < j; i++) {
      const double avg = 0.5 * (T[i + (size_t) j * ell] + T[j + (size_t) i * ell]);
      T[i + (size_t) j * ell] = avg; T[j + (size_t) i * ell] = avg;
    }
  return orc_syev (ell, T, ell, evals);
}

/* glr_common_impl.h:415-608; with opts.backend == 2 the orthonormalization
 * of glr_scalapack_impl.h:573-586 (BCGS2, 64-column panels) and no
 * residual sweep (that backend's check compares two eigensolvers: this
 * restatement has one, and leaves resid_max at 0) */
static int
build_replicated (orc_glr_t *g, orc_glr_report_t *rep)
{
  const int n = g->n, ell = g->opts.ell;
  double *Y = xcalloc ((size_t) n * ell, sizeof (double));
  double *Q = xcalloc ((size_t) n * ell, sizeof (double));
  double *G = xcalloc ((size_t) ell * ell, sizeof (double));
  double *A = xcalloc ((size_t) ell * ell, sizeof (double));
  double *T = xcalloc ((size_t) ell * ell, sizeof (double));
  double *evals = xcalloc ((size_t) ell, sizeof (double));
  int *order = NULL, rc = 0;
  const int bcgs_build = g->opts.backend == 2;
  double t0;

  memset (rep, 0, sizeof (*rep));
  orc_randn_fill (g->opts.seed, 0, 0, n, ell, Q, n);
  t0 = now_s ();
  orc_glr_apply_F (g, ell, Q, n, Y, n);
  rep->t_operator += now_s () - t0;
  for (int p = 0; p < g->opts.q_power; p++) {
    t0 = now_s ();
    if (bcgs_build ? bcgs2 (n, Y, ell, 64, 0) : cholqr2 (n, ell, Y, G)) { rc = -1; goto out; }
    rep->t_dense += now_s () - t0;
    memcpy (Q, Y, sizeof (double) * (size_t) n * ell);
    t0 = now_s ();
    orc_glr_apply_F (g, ell, Q, n, Y, n);
    rep->t_operator += now_s () - t0;
  }
  t0 = now_s ();
  if (bcgs_build ? bcgs2 (n, Y, ell, 64, 0) : cholqr2 (n, ell, Y, G)) { rc = -1; goto out; }
  rep->t_dense += now_s () - t0;
  memcpy (Q, Y, sizeof (double) * (size_t) n * ell);
  t0 = now_s ();
  orc_glr_apply_F (g, ell, Q, n, Y, n);
  rep->t_operator += now_s () - t0;
  orc_gemm_tn (ell, ell, n, Q, n, Y, n, A, ell);
  t0 = now_s ();
  if (eig_of_raw (ell, A, T, evals)) { rc = -2; goto out; }
  rep->t_dense += now_s () - t0;
  select_and_form_U (g, ell, Q, T, evals, rep, &order);
  if (!bcgs_build) residual_check (g, evals, order, rep, &rep->t_operator);
  free (g->Q); free (g->Araw);
  g->Q = Q; Q = NULL;
  g->Araw = A; A = NULL;
  g->ell = ell;
out:
  free (Y); free (Q); free (G); free (A); free (T); free (evals); free (order);
  return rc;
}

orc_glr_t *
orc_glr_build (int n, const int *brp, const int *bci, const double *bv,
               const double *mass, orc_zs_t *zs, const orc_glr_opts_t *opts,
               orc_glr_report_t *rep)
{
  orc_glr_t *g;
  orc_glr_report_t lrep;
  if (opts->ell <= 0 || opts->ell > n) return NULL;   /* :651-654 */
  g = xcalloc (1, sizeof (*g));
  g->n = n; g->brp = brp; g->bci = bci; g->bv = bv; g->mass = mass;
  g->zs = zs; g->opts = *opts;
  g->msqrt = xcalloc ((size_t) n, sizeof (double));
  g->minvsqrt = xcalloc ((size_t) n, sizeof (double));
  for (int i = 0; i < n; i++) {                 /* glr_common_impl.h:123-128 */
    g->msqrt[i] = sqrt (fabs (mass[i]));
    g->minvsqrt[i] = 1. / g->msqrt[i];
  }
  g->w1 = xcalloc ((size_t) n, sizeof (double));
  g->w2 = xcalloc ((size_t) n, sizeof (double));
  if (build_replicated (g, &lrep)) { orc_glr_free (g); return NULL; }
  if (rep) *rep = lrep;
  return g;
}

/* incremental extension: glr_scalapack_impl.h:689-915 with the raw A
 * kept replicated (comm size 1) and U materialized as the replicated
 * backend does; incremental = 0 is the replicated rebuild
 * (glr_common_impl.h:711-715).                                          */
int
orc_glr_extend (orc_glr_t *g, int k_new, int incremental, int panel,
                orc_glr_report_t *rep)
{
  const int n = g->n;
  orc_glr_report_t lrep;
  if (k_new <= 0 || g->opts.ell + k_new > n) return -1;   /* :699-704 */
  if (!incremental) {
    g->opts.ell += k_new;
    if (build_replicated (g, &lrep)) return -2;
    if (rep) *rep = lrep;
    return 0;
  }
  {
    const int ell1 = g->ell, ell2 = g->ell + k_new;
    double *Qb = xcalloc ((size_t) n * ell2, sizeof (double));
    double *Qn = xcalloc ((size_t) n * k_new, sizeof (double));
    double *Yn = xcalloc ((size_t) n * k_new, sizeof (double));
    double *Cn = xcalloc ((size_t) ell2 * k_new, sizeof (double));
    double *A2 = xcalloc ((size_t) ell2 * ell2, sizeof (double));
    double *T = xcalloc ((size_t) ell2 * ell2, sizeof (double));
    double *evals = xcalloc ((size_t) ell2, sizeof (double));
    int *order = NULL;
    double t0;

    memset (&lrep, 0, sizeof (lrep));
    memcpy (Qb, g->Q, sizeof (double) * (size_t) n * ell1);
    orc_randn_fill (g->opts.seed, 0, ell1, n, k_new, Qn, n);
    t0 = now_s ();
    orc_glr_apply_F (g, k_new, Qn, n, Yn, n);
    lrep.t_operator += now_s () - t0;
    for (int p = 0; p < g->opts.q_power; p++) {
      t0 = now_s ();
      if (bcgs2 (n, Yn, k_new, panel, 0)) return -3;
      lrep.t_dense += now_s () - t0;
      memcpy (Qn, Yn, sizeof (double) * (size_t) n * k_new);
      t0 = now_s ();
      orc_glr_apply_F (g, k_new, Qn, n, Yn, n);
      lrep.t_operator += now_s () - t0;
    }
    memcpy (Qb + (size_t) ell1 * n, Yn, sizeof (double) * (size_t) n * k_new);
    t0 = now_s ();
    if (bcgs2 (n, Qb, ell2, panel, ell1)) return -3;
    lrep.t_dense += now_s () - t0;
    memcpy (Qn, Qb + (size_t) ell1 * n, sizeof (double) * (size_t) n * k_new);
    t0 = now_s ();
    orc_glr_apply_F (g, k_new, Qn, n, Yn, n);
    lrep.t_operator += now_s () - t0;
    orc_gemm_tn (ell2, k_new, n, Qb, n, Yn, n, Cn, ell2);
    /* bordered raw A (:819-840): old block kept; new rows of old columns
     * by symmetry; new columns from Cn                                  */
    for (int j = 0; j < ell1; j++) {
      for (int i = 0; i < ell1; i++)
        A2[i + (size_t) j * ell2] = g->Araw[i + (size_t) j * ell1];
      for (int i = ell1; i < ell2; i++)
        A2[i + (size_t) j * ell2] = Cn[j + (size_t) (i - ell1) * ell2];
    }
    for (int j = ell1; j < ell2; j++)
      for (int i = 0; i < ell2; i++)
        A2[i + (size_t) j * ell2] = Cn[i + (size_t) (j - ell1) * ell2];
    t0 = now_s ();
    if (eig_of_raw (ell2, A2, T, evals)) return -4;
    lrep.t_dense += now_s () - t0;
    g->opts.ell = ell2;
    select_and_form_U (g, ell2, Qb, T, evals, &lrep, &order);
    residual_check (g, evals, order, &lrep, &lrep.t_operator);
    free (g->Q); free (g->Araw);
    g->Q = Qb; g->Araw = A2; g->ell = ell2;
    free (Qn); free (Yn); free (Cn); free (T); free (evals); free (order);
    if (rep) *rep = lrep;
  }
  return 0;
}

void
orc_glr_free (orc_glr_t *g)
{
  if (!g) return;
  free (g->msqrt); free (g->minvsqrt); free (g->U); free (g->lam);
  free (g->Q); free (g->Araw); free (g->w1); free (g->w2);
  free (g);
}

int orc_glr_kept (const orc_glr_t *g) { return g->kept; }
const double *orc_glr_lam (const orc_glr_t *g) { return g->lam; }
const double *orc_glr_U (const orc_glr_t *g) { return g->U; }

/* ---- downstream: glr_common_impl.h:947-1104 ---------------------------- */

/* out = f_c in + U diag(f_lam - f_c) U^T in (:948-966) */
static void
filter_core (orc_glr_t *g, const double *f_lam, double f_c, const double *in,
             double *out)
{
  const int n = g->n, k = g->kept;
  if (k == 0) {
    for (int i = 0; i < n; i++) out[i] = in[i] * f_c;
    return;
  }
  {
    double *w = xcalloc ((size_t) k, sizeof (double));
    for (int j = 0; j < k; j++) {
      const double *u = g->U + (size_t) j * n;
      double s = 0.;
      for (int i = 0; i < n; i++) s += u[i] * in[i];
      w[j] = s * (f_lam[j] - f_c);
    }
    for (int i = 0; i < n; i++) out[i] = in[i] * f_c;
    for (int j = 0; j < k; j++) {
      const double *u = g->U + (size_t) j * n;
      for (int i = 0; i < n; i++) out[i] += u[i] * w[j];
    }
    free (w);
  }
}

int
orc_glr_filter (orc_glr_t *g, const double *f_lam, double f_c,
                const double *in, double *out)
{
  filter_core (g, f_lam, f_c, in, out);
  return 0;
}

int
orc_glr_pow (orc_glr_t *g, double c, double p, const double *in, double *out)
{
  double *f;
  if (!(c > 0.)) return -1;                      /* :980-982, floor = 0 */
  f = xcalloc ((size_t) MAX (g->kept, 1), sizeof (double));
  for (int i = 0; i < g->kept; i++) f[i] = pow (g->lam[i] + c, p);
  filter_core (g, f, pow (c, p), in, out);
  free (f);
  return 0;
}

double
orc_glr_logdet (orc_glr_t *g, double c)
{
  double s = 0.;
  for (int i = 0; i < g->kept; i++) s += log (g->lam[i] + c) - log (c);
  return s;
}

int
orc_glr_solve (orc_glr_t *g, double c, const double *rhs, double *x)
{
  const int n = g->n;
  double *f;
  if (!(c > 0.)) return -1;                      /* :1013-1015 */
  f = xcalloc ((size_t) MAX (g->kept, 1), sizeof (double));
  orc_zs_solve_vec (g->zs, rhs, g->w1);
  for (int i = 0; i < n; i++) g->w1[i] *= g->msqrt[i];
  for (int i = 0; i < g->kept; i++) f[i] = 1. / (g->lam[i] + c);
  filter_core (g, f, 1. / c, g->w1, g->w2);
  for (int i = 0; i < n; i++) g->w2[i] *= g->msqrt[i];
  orc_zs_solve_vec (g->zs, g->w2, x);
  free (f);
  return 0;
}

int
orc_glr_apply (orc_glr_t *g, double c, const double *v, double *y)
{
  /* exact: y = B v + c Z M^{-1} Z^T v  (:1043-1047, prior apply :236-244) */
  const int n = g->n;
  orc_csr_mm (n, g->brp, g->bci, g->bv, 1, v, n, y, n);
  orc_zs_apply (g->zs, 1, v, n, g->w1, n);
  for (int i = 0; i < n; i++) g->w1[i] = g->w1[i] / g->mass[i];
  orc_zs_apply (g->zs, 1, g->w1, n, g->w2, n);
  for (int i = 0; i < n; i++) y[i] += c * g->w2[i];
  return 0;
}

int
orc_glr_factor (orc_glr_t *g, double c, int which, const double *in,
                double *out)
{
  const int n = g->n, k = g->kept;
  const double half = 0.5;
  double *f;
  if (!(c > 0.)) return -1;
  f = xcalloc ((size_t) MAX (k, 1), sizeof (double));
  switch (which) {
  case 0:   /* G = Z M^{-1/2} S */
    for (int i = 0; i < k; i++) f[i] = pow (g->lam[i] + c, half);
    filter_core (g, f, sqrt (c), in, g->w1);
    for (int i = 0; i < n; i++) g->w1[i] *= g->minvsqrt[i];
    orc_zs_apply (g->zs, 1, g->w1, n, out, n);
    break;
  case 1:   /* GT = S M^{-1/2} Z^T */
    orc_zs_apply (g->zs, 1, in, n, g->w1, n);
    for (int i = 0; i < n; i++) g->w1[i] *= g->minvsqrt[i];
    for (int i = 0; i < k; i++) f[i] = pow (g->lam[i] + c, half);
    filter_core (g, f, sqrt (c), g->w1, out);
    break;
  case 2:   /* GINV = S^{-1} M^{1/2} Z^{-1} */
    orc_zs_solve_vec (g->zs, in, g->w1);
    for (int i = 0; i < n; i++) g->w1[i] *= g->msqrt[i];
    for (int i = 0; i < k; i++) f[i] = 1. / pow (g->lam[i] + c, half);
    filter_core (g, f, 1. / sqrt (c), g->w1, out);
    break;
  case 3:   /* GINVT = Z^{-T} M^{1/2} S^{-1} */
    for (int i = 0; i < k; i++) f[i] = 1. / pow (g->lam[i] + c, half);
    filter_core (g, f, 1. / sqrt (c), in, g->w1);
    for (int i = 0; i < n; i++) g->w1[i] *= g->msqrt[i];
    orc_zs_solve_vec (g->zs, g->w1, out);
    break;
  default:
    free (f);
    return -2;
  }
  free (f);
  return 0;
}

int
orc_glr_sample (orc_glr_t *g, double c, const double *xi, double *x)
{
  return orc_glr_factor (g, c, 3, xi, x);        /* :1099-1104 */
}
