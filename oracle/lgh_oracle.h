/* lgh_oracle.h — FP64 CPU restatement of lgpsf-hessian's GLR stage and its
 * downstream operations, used ONLY as the parity reference and the timed CPU
 * baseline.  TEST INFRASTRUCTURE: nothing in the product library links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may.
 *
 * Every function follows a reference function step for step (file:line
 * given at each definition; paths relative to the reference repository).
 * All dense blocks are column-major with an explicit leading dimension,
 * exactly as the reference's MATDENSE row slabs at communicator size 1.
 *
 * Pinning status of each part (what checks it besides restating it):
 *   randn_at         — statistical moments + partition/extension invariants
 *                      + an independent Python-integer restatement; the
 *                      reference ships no stored draws (parity pinned only
 *                      to the published construction).
 *   dense kernels    — numpy/LAPACK golden fixtures (eigvalsh, cholesky).
 *   GLR build/ops    — tests/test_glr.cpp closed-form problem vs numpy
 *                      dense-oracle eigenvalues; every identity gate of
 *                      tests/test_glr.cpp:221-316 at its tolerance.
 *   Z-solve tiers    — tests/test_prior_mat.cpp 24x24 Laplacian + 0.5 I vs
 *                      a scipy direct solve (tolerances of :144-190).
 *   Chebyshev count  — closed form of zsolve_impl.h:93-102.
 */
#ifndef LGH_ORACLE_H
#define LGH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashed Gaussian kit (glr_common_impl.h:32-73) ------------------- */
double orc_randn_at (uint64_t seed, int64_t i, int64_t j);
/* out[(r) + (c)*ld] = randn_at(seed, row0 + r, col0 + c) */
void   orc_randn_fill (uint64_t seed, int64_t row0, int64_t col0,
                       int64_t nr, int64_t nc, double *out, int64_t ld);

/* ---- CSR --------------------------------------------------------------- */
/* Y = A X on column-major blocks (PETSc MatMatMult / per-column MatMult
 * row order: each row sums its nonzeros in stored order).              */
/* host threads for the column-parallel sweeps (block Z-solves and sparse
 * products; results do not depend on the count); default 1 */
void orc_set_threads (int nt);
int  orc_get_threads (void);

void orc_csr_mm (int n, const int *rp, const int *ci, const double *v,
                 int ncols, const double *X, int ldx, double *Y, int ldy);

/* ---- multigrid hierarchy (zsolve_impl.h:38-65) -------------------------
 * Levels l = 0 (coarsest) .. nl-1 (finest), PCMG numbering.  Level l > 0
 * carries A_l, P_l (maps level l-1 -> l), dinv, smoother interval and
 * degree; level 0 is solved by a dense LU of A_0.  All arrays are copied. */
typedef struct orc_mg orc_mg_t;
orc_mg_t *orc_mg_create (int nl, int tile);
int  orc_mg_set_level (orc_mg_t *h, int l, int n, const int *rp,
                       const int *ci, const double *v, int nc_prev,
                       const int *prp, const int *pci, const double *pv,
                       double slo, double shi, int nu);
void orc_mg_free (orc_mg_t *h);
/* 1: V-cycle (default), 2: W-cycle */
void orc_mg_set_cycle (orc_mg_t *h, int gamma);
/* X = V-cycle(B) (zero initial guess), ncols columns                    */
void orc_mg_vcycle (orc_mg_t *h, int ncols, const double *B, int ldb,
                    double *X, int ldx);

/* ---- Z-solvers --------------------------------------------------------- */
typedef struct orc_zs orc_zs_t;
/* exact diagonal Z (the callbacks prior of tests/test_glr.cpp:67-85)     */
orc_zs_t *orc_zs_diag (int n, const double *zdiag);
/* assembled Z; vector solves by PCG (Jacobi, or V-cycle when h != NULL)
 * to rtol; block solves per `mode`:
 *   0 — per-column vector solves (zsolve_impl.h:685-715 mode < 2)
 *   3 — blocked Chebyshev around the V-cycle on h (zsolve_impl.h:599-644)
 *   4 — blocked Chebyshev with the Jacobi row-scaling PC (the mode-3
 *       level smoother used as the outer solver)
 * lmin/lmax/nit: the margined outer interval and iteration count.       */
orc_zs_t *orc_zs_mat (int n, const int *rp, const int *ci, const double *v,
                      orc_mg_t *h, int mode, double lmin, double lmax,
                      int nit, double ksp_rtol);
void orc_zs_free (orc_zs_t *zs);
void orc_zs_apply (orc_zs_t *zs, int ncols, const double *X, int ldx,
                   double *Y, int ldy);
void orc_zs_solve_vec (orc_zs_t *zs, const double *b, double *x);
void orc_zs_solve_block (orc_zs_t *zs, int ncols, const double *X, int ldx,
                         double *Y, int ldy);
/* Chebyshev on blocks, arbitrary PC: zsolve_impl.h:108-141 (exposed for
 * operator-level parity: X = p(P A) P B with zero initial guess).        */
void orc_cheb_jacobi (int n, const int *rp, const int *ci, const double *v,
                      const double *dinv, double lmin, double lmax, int nit,
                      int ncols, const double *B, int ldb, double *X,
                      int ldx);
void orc_cheb_mg (orc_mg_t *h, double lmin, double lmax, int nit, int ncols,
                  const double *B, int ldb, double *X, int ldx);
int  orc_cheb_count (double lmin, double lmax, double tol);
/* PCG on one vector with Lanczos extreme-eigenvalue estimates of the
 * preconditioned operator (KSPComputeExtremeSingularValues semantics).
 * pc: 0 = Jacobi, 1 = V-cycle on h.  Returns iterations.                 */
int  orc_cg (int n, const int *rp, const int *ci, const double *v, int pc,
             orc_mg_t *h, const double *b, double *x, double rtol,
             int maxit, double *emin, double *emax);

/* ---- dense (the f77_decls.h:20-24 kernels, own arithmetic) ------------ */
int  orc_potrf_l (int n, double *G, int ld);            /* G = L L^T     */
int  orc_potrf_u (int n, double *G, int ld);            /* G = R^T R     */
void orc_trsm_rltn (int m, int n, const double *L, int ldl, double *X,
                    int ldx);                           /* X <- X L^{-T} */
void orc_trsm_runn (int m, int n, const double *R, int ldr, double *X,
                    int ldx);                           /* X <- X R^{-1} */
/* symmetric eigensolver: ascending w, orthonormal vectors in A (uses the
 * upper and lower triangle as given; callers pass symmetric A).          */
int  orc_syev (int n, double *A, int lda, double *w);
void orc_gemm_tn (int m, int n, int k, const double *A, int lda,
                  const double *B, int ldb, double *C, int ldc);
void orc_gemm_nn (int m, int n, int k, double alpha, const double *A,
                  int lda, const double *B, int ldb, double beta, double *C,
                  int ldc);

/* ---- GLR (glr.h:68-169 semantics) --------------------------------------- */
typedef struct {
  int      ell, q_power;
  double   trunc_abs, trunc_rel;
  int      treat;             /* 0 FLIP, 1 RELU, 2 NONE                    */
  uint64_t seed;
  int      check;
  int      backend;           /* 2: the at-scale backend's build at one
                                 process (BCGS2 in 64-column panels, no
                                 residual sweep); otherwise replicated     */
} orc_glr_opts_t;

typedef struct {
  int    kept;
  double lam_abs_max, lam_raw_min, lam_raw_max, next_abs;
  int    n_negative_raw;
  double resid_max, t_operator, t_dense;
} orc_glr_report_t;

typedef struct orc_glr orc_glr_t;

/* B (CSR, n x n) and the prior (mass + Z-solver) are borrowed.          */
orc_glr_t *orc_glr_build (int n, const int *brp, const int *bci,
                          const double *bv, const double *mass,
                          orc_zs_t *zs, const orc_glr_opts_t *opts,
                          orc_glr_report_t *rep);
int  orc_glr_extend (orc_glr_t *g, int k_new, int incremental, int panel,
                     orc_glr_report_t *rep);
void orc_glr_free (orc_glr_t *g);
int  orc_glr_kept (const orc_glr_t *g);
const double *orc_glr_lam (const orc_glr_t *g);
const double *orc_glr_U (const orc_glr_t *g);           /* n x max(k,1)  */
/* one F sweep (exposed for operator-level parity): Y = F X             */
void orc_glr_apply_F (orc_glr_t *g, int ncols, const double *X, int ldx,
                      double *Y, int ldy);
int    orc_glr_solve  (orc_glr_t *g, double c, const double *rhs, double *x);
int    orc_glr_apply  (orc_glr_t *g, double c, const double *v, double *y);
int    orc_glr_factor (orc_glr_t *g, double c, int which, const double *in,
                       double *out);
int    orc_glr_sample (orc_glr_t *g, double c, const double *xi, double *x);
int    orc_glr_pow    (orc_glr_t *g, double c, double p, const double *in,
                       double *out);
int    orc_glr_filter (orc_glr_t *g, const double *f_lam, double f_c,
                       const double *in, double *out);
double orc_glr_logdet (orc_glr_t *g, double c);

#ifdef __cplusplus
}
#endif
#endif
