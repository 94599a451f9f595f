/* test_glr_api.cpp — a C++ consumer of the unchanged public API
 * (lgpsf_hessian.h) on the closed-form problem of tests/test_glr.cpp:43-65:
 * N = 120, rank-10 B (dense Mat), diagonal Z through the Vec callbacks.
 * Gates (tolerances of tests/test_glr.cpp:221-316): kept == 10, spectrum
 * vs the dense eigenvalues passed in argv[1] (|lambda|-descending, one per
 * line), next_abs, residual gate, logdet, solve o apply, G G^T, GINVT GT,
 * sample == GINVT, pow o pow^-1, filter == pow(-1/2), extend == one-shot.
 * Links liblgh.so; prints PASS and exits 0 when every gate holds. */
#include <lgpsf_hessian/lgpsf_hessian.h>

#include <math.h>
#include <stdio.h>

#define N 120
#define RANK 10

static int nfail = 0;

static void
gate (int ok, const char *what, double v)
{
  if (!ok) { nfail++; printf ("FAIL: %s (%.3e)\n", what, v); }
}

static double a_k (int k, int i) { return sin (0.1 * (i + 1) * (k + 3)) + 0.2 * cos (0.37 * i * k + k); }
static double zdiag (int i) { return 1.0 + 0.01 * i; }
static double mlump (int i) { return 0.5 + 0.1 * (i % 7); }
static double probe (int i) { return cos (0.05 * i + 1.0); }

static double
b_ij (int i, int j, int spd)
{
  double s = 0.;
  for (int k = 0; k < RANK; k++) {
    double sk = 2.0 / (k + 1.0);
    if (!spd && (k % 2)) sk = -sk;
    s += sk * a_k (k, i) * a_k (k, j);
  }
  return s;
}

static Vec g_z;

static void cb_apply (Vec x, Vec y, void *) { VecPointwiseMult (y, x, g_z); }
static void cb_solve (Vec x, Vec y, void *) { VecPointwiseDivide (y, x, g_z); }

static void
fill (Vec v, double (*f) (int))
{
  PetscScalar *a;
  VecGetArray (v, &a);
  for (int i = 0; i < N; i++) a[i] = f (i);
  VecRestoreArray (v, &a);
}

static double
rdiff (Vec a, Vec b)
{
  Vec d;
  double nd, nb;
  VecDuplicate (a, &d);
  VecCopy (a, d);
  VecAXPY (d, -1.0, b);
  VecNorm (d, NORM_2, &nd);
  VecNorm (b, NORM_2, &nb);
  VecDestroy (&d);
  return nd / nb;
}

static int
scenario (int spd, const double *dense)
{
  const double c = 0.7;
  Mat B;
  Vec mass, v, y, x, s2;
  lgh_prior_t *prior;
  lgh_glr_t *glr, *glr2;
  lgh_glr_report_t rep, rep2;
  lgh_prior_callbacks_t cb = { cb_apply, NULL, cb_solve, NULL, NULL, NULL, NULL };

  PetscCall (MatCreateDense (PETSC_COMM_WORLD, PETSC_DECIDE, PETSC_DECIDE, N, N, NULL, &B));
  {
    PetscScalar *ba;
    PetscInt lda;
    PetscCall (MatDenseGetLDA (B, &lda));
    PetscCall (MatDenseGetArrayWrite (B, &ba));
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) ba[i + j * lda] = b_ij (i, j, spd);
    PetscCall (MatDenseRestoreArrayWrite (B, &ba));
  }
  PetscCall (MatAssemblyBegin (B, MAT_FINAL_ASSEMBLY));
  PetscCall (MatAssemblyEnd (B, MAT_FINAL_ASSEMBLY));
  PetscCall (MatCreateVecs (B, &mass, NULL));
  VecDuplicate (mass, &g_z); VecDuplicate (mass, &v); VecDuplicate (mass, &y);
  VecDuplicate (mass, &x); VecDuplicate (mass, &s2);
  fill (mass, mlump); fill (g_z, zdiag); fill (v, probe);
  PetscCall (lgh_prior_create_callbacks (mass, &cb, &prior));

  lgh_glr_opts_t go = lgh_glr_opts_default ();
  go.ell = 40; go.q_power = 1; go.trunc_rel = 1e-8; go.nb = 8; go.panel = 16;
  go.check = 1;
  PetscCall (lgh_glr_compute (B, prior, &go, &glr, &rep));

  gate (rep.kept == RANK, "kept == rank(B)", rep.kept);
  {
    int nk; const double *lam; double worst = 0.;
    lgh_glr_eigs (glr, &nk, &lam);
    for (int i = 0; i < nk && i < RANK; i++)
      worst = fmax (worst, fabs (lam[i] - fabs (dense[i])) / fabs (dense[i]));
    gate (worst < 1e-8, "spectrum vs dense", worst);
  }
  gate (rep.next_abs < 1e-6 * rep.lam_abs_max, "next_abs", rep.next_abs);
  gate (rep.resid_max < 1e-7 * rep.lam_abs_max, "residual gate", rep.resid_max);
  if (!spd) gate (rep.n_negative_raw >= RANK / 2, "negatives", rep.n_negative_raw);
  {
    double ref = 0.;
    for (int i = 0; i < RANK; i++) ref += log (fabs (dense[i]) + c) - log (c);
    gate (fabs (lgh_glr_logdet (glr, c) - ref) < 1e-9 * fabs (ref), "logdet",
          lgh_glr_logdet (glr, c) - ref);
  }
  if (spd) {
    double rd;
    PetscCall (lgh_glr_apply (glr, c, v, y));
    PetscCall (lgh_glr_solve (glr, c, y, x));
    rd = rdiff (x, v); gate (rd < 1e-8, "solve(apply(v)) == v", rd);
    PetscCall (lgh_glr_factor (glr, c, LGH_FACTOR_GT, v, x));
    PetscCall (lgh_glr_factor (glr, c, LGH_FACTOR_G, x, s2));
    rd = rdiff (s2, y); gate (rd < 1e-8, "G(GT(v)) == apply(v)", rd);
    PetscCall (lgh_glr_factor (glr, c, LGH_FACTOR_GT, v, y));
    PetscCall (lgh_glr_factor (glr, c, LGH_FACTOR_GINVT, y, x));
    rd = rdiff (x, v); gate (rd < 1e-8, "GINVT(GT(v)) == v", rd);
    PetscCall (lgh_glr_sample (glr, c, y, s2));
    rd = rdiff (s2, x); gate (rd < 1e-14, "sample == GINVT", rd);
    PetscCall (lgh_glr_pow (glr, c, -1.0, v, y));
    PetscCall (lgh_glr_pow (glr, c, 1.0, y, x));
    rd = rdiff (x, v); gate (rd < 1e-10, "pow(+1)(pow(-1)(v)) == v", rd);
    {
      int nk; const double *lam; double fl[N];
      lgh_glr_eigs (glr, &nk, &lam);
      for (int i = 0; i < nk; i++) fl[i] = 1.0 / sqrt (lam[i] + c);
      PetscCall (lgh_glr_filter (glr, fl, 1.0 / sqrt (c), v, y));
      PetscCall (lgh_glr_pow (glr, c, -0.5, v, x));
      rd = rdiff (y, x); gate (rd < 1e-12, "filter == pow(-1/2)", rd);
    }
  }
  {
    lgh_glr_opts_t go2 = go;
    int nk1, nk2; const double *l1, *l2; double worst = 0.;
    go2.ell = 30;
    PetscCall (lgh_glr_compute (B, prior, &go2, &glr2, &rep2));
    PetscCall (lgh_glr_extend (glr2, 10, &rep2));
    lgh_glr_eigs (glr, &nk1, &l1);
    lgh_glr_eigs (glr2, &nk2, &l2);
    gate (nk1 == nk2, "extend: kept", nk2);
    for (int i = 0; i < nk1 && i < nk2; i++) worst = fmax (worst, fabs (l1[i] - l2[i]) / fabs (l1[i]));
    gate (worst < 1e-10, "extend: spectrum", worst);
    lgh_glr_destroy (glr2);
  }
  /* refusals (glr_common_impl.h:651-654, 699-704, 1013-1015) */
  gate (lgh_glr_solve (glr, 0.0, v, x) == PETSC_ERR_ARG_OUTOFRANGE, "c <= floor refused", 0);
  gate (lgh_glr_extend (glr, 0, NULL) == PETSC_ERR_ARG_OUTOFRANGE, "k_new <= 0 refused", 0);

  lgh_glr_destroy (glr);
  lgh_prior_destroy (prior);
  VecDestroy (&mass); VecDestroy (&g_z); VecDestroy (&v); VecDestroy (&y);
  VecDestroy (&x); VecDestroy (&s2);
  MatDestroy (&B);
  return 0;
}

int
main (int argc, char **argv)
{
  double dense[2][N];
  if (argc < 3) { fprintf (stderr, "usage: %s eigs_spd.txt eigs_indef.txt\n", argv[0]); return 2; }
  for (int s = 0; s < 2; s++) {
    FILE *f = fopen (argv[1 + s], "r");
    if (!f) { fprintf (stderr, "cannot open %s\n", argv[1 + s]); return 2; }
    for (int i = 0; i < N; i++)
      if (fscanf (f, "%lf", &dense[s][i]) != 1) { fclose (f); return 2; }
    fclose (f);
  }
  PetscCall (PetscInitialize (&argc, &argv, NULL, NULL));
  if (scenario (1, dense[0])) return 3;
  if (scenario (0, dense[1])) return 3;
  if (nfail == 0) PetscPrintf (PETSC_COMM_WORLD, "PASS test_glr_api\n");
  PetscCall (PetscFinalize ());
  return nfail == 0 ? 0 : 1;
}
