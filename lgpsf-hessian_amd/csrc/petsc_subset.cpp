// petsc_subset.cpp — sequential host Vec / Mat (AIJ and DENSE) with the
// PETSc names the GLR API boundary uses, plus the device mirror of a Mat
// (CSR uploaded once per precision and reused until the Mat changes, as
// PETSc's GPU matrix types keep an offloaded copy).
#include "petsc_impl.hpp"
#include "lgh_dev.hpp"

#include <chrono>
#include <cstdarg>
#include <map>
#include <string>

static thread_local std::string g_perr;

extern "C" {

PetscErrorCode
lgh_petsc_error (PetscErrorCode code, const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start (ap, fmt);
  vsnprintf (buf, sizeof buf, fmt, ap);
  va_end (ap);
  g_perr = buf;
  return code;
}

const char *lgh_petsc_last_error (void) { return g_perr.c_str (); }

PetscErrorCode PetscInitialize (int *, char ***, const char *, const char *)
{ return 0; }
PetscErrorCode PetscFinalize (void) { return 0; }

PetscErrorCode
PetscPrintf (MPI_Comm, const char *fmt, ...)
{
  va_list ap;
  va_start (ap, fmt);
  vprintf (fmt, ap);
  va_end (ap);
  fflush (stdout);
  return 0;
}

double
MPI_Wtime (void)
{
  using namespace std::chrono;
  return duration<double> (steady_clock::now ().time_since_epoch ()).count ();
}

PetscErrorCode
PetscObjectReference (PetscObject obj)
{
  if (obj) ((lgh_obj_head *) obj)->refct++;
  return 0;
}

PetscErrorCode
PetscObjectGetComm (PetscObject, MPI_Comm *comm)
{
  *comm = PETSC_COMM_WORLD;
  return 0;
}

// ---- Vec -------------------------------------------------------------------

static PetscErrorCode
vec_new (PetscInt n, Vec *v)
{
  PetscCheck (n >= 0, 0, PETSC_ERR_ARG_OUTOFRANGE, "VecCreate: negative size");
  Vec x = new _p_Vec ();
  x->n = n;
  x->own.assign ((size_t) n, 0.0);
  x->a = x->own.data ();
  *v = x;
  return 0;
}

PetscErrorCode VecCreateSeq (MPI_Comm, PetscInt n, Vec *v) { return vec_new (n, v); }

PetscErrorCode
VecCreateMPI (MPI_Comm, PetscInt n, PetscInt N, Vec *v)
{
  return vec_new (n >= 0 ? n : N, v);
}

PetscErrorCode
VecCreateSeqWithArray (MPI_Comm, PetscInt, PetscInt n, const PetscScalar *a,
                       Vec *v)
{
  Vec x = new _p_Vec ();
  x->n = n;
  x->a = (double *) a;          // borrowed, as PETSc does
  *v = x;
  return 0;
}

PetscErrorCode VecDuplicate (Vec x, Vec *y) { return vec_new (x->n, y); }

PetscErrorCode
VecDestroy (Vec *v)
{
  if (!v || !*v) return 0;
  if (--(*v)->refct <= 0) delete *v;
  *v = NULL;
  return 0;
}

PetscErrorCode
VecCopy (Vec x, Vec y)
{
  PetscCheck (x->n == y->n, 0, PETSC_ERR_ARG_SIZ, "VecCopy: size mismatch");
  if (x->a != y->a) memcpy (y->a, x->a, sizeof (double) * x->n);
  return 0;
}

PetscErrorCode VecSet (Vec x, PetscScalar a)
{ for (int i = 0; i < x->n; i++) x->a[i] = a; return 0; }
PetscErrorCode VecZeroEntries (Vec x) { return VecSet (x, 0.0); }
PetscErrorCode VecScale (Vec x, PetscScalar a)
{ for (int i = 0; i < x->n; i++) x->a[i] *= a; return 0; }

PetscErrorCode
VecAXPY (Vec y, PetscScalar a, Vec x)
{
  PetscCheck (x->n == y->n, 0, PETSC_ERR_ARG_SIZ, "VecAXPY: size mismatch");
  for (int i = 0; i < x->n; i++) y->a[i] += a * x->a[i];
  return 0;
}

PetscErrorCode
VecNorm (Vec x, NormType t, PetscReal *nrm)
{
  double s = 0.0;
  if (t == NORM_2) { for (int i = 0; i < x->n; i++) s += x->a[i] * x->a[i]; s = sqrt (s); }
  else if (t == NORM_1) { for (int i = 0; i < x->n; i++) s += fabs (x->a[i]); }
  else { for (int i = 0; i < x->n; i++) s = PetscMax (s, fabs (x->a[i])); }
  *nrm = s;
  return 0;
}

PetscErrorCode
VecDot (Vec x, Vec y, PetscScalar *d)
{
  double s = 0.0;
  for (int i = 0; i < x->n; i++) s += x->a[i] * y->a[i];
  *d = s;
  return 0;
}

PetscErrorCode
VecPointwiseMult (Vec w, Vec x, Vec y)
{
  for (int i = 0; i < w->n; i++) w->a[i] = x->a[i] * y->a[i];
  return 0;
}

PetscErrorCode
VecPointwiseDivide (Vec w, Vec x, Vec y)
{
  for (int i = 0; i < w->n; i++) w->a[i] = x->a[i] / y->a[i];
  return 0;
}

PetscErrorCode VecSqrtAbs (Vec x)
{ for (int i = 0; i < x->n; i++) x->a[i] = sqrt (fabs (x->a[i])); return 0; }
PetscErrorCode VecReciprocal (Vec x)
{ for (int i = 0; i < x->n; i++) if (x->a[i] != 0.0) x->a[i] = 1.0 / x->a[i]; return 0; }
PetscErrorCode VecGetSize (Vec x, PetscInt *n) { *n = x->n; return 0; }
PetscErrorCode VecGetLocalSize (Vec x, PetscInt *n) { *n = x->n; return 0; }
PetscErrorCode VecGetOwnershipRange (Vec x, PetscInt *lo, PetscInt *hi)
{ if (lo) *lo = 0; if (hi) *hi = x->n; return 0; }
PetscErrorCode VecGetArray (Vec x, PetscScalar **a) { *a = x->a; return 0; }
PetscErrorCode VecRestoreArray (Vec, PetscScalar **a) { *a = NULL; return 0; }
PetscErrorCode VecGetArrayRead (Vec x, const PetscScalar **a) { *a = x->a; return 0; }
PetscErrorCode VecRestoreArrayRead (Vec, const PetscScalar **a) { *a = NULL; return 0; }
PetscErrorCode VecGetArrayWrite (Vec x, PetscScalar **a) { *a = x->a; return 0; }
PetscErrorCode VecRestoreArrayWrite (Vec, PetscScalar **a) { *a = NULL; return 0; }

// ---- Mat -------------------------------------------------------------------

static Mat
mat_new (int type, PetscInt m, PetscInt n)
{
  Mat A = new _p_Mat ();
  A->type = type; A->m = m; A->n = n;
  return A;
}

PetscErrorCode
MatCreateAIJ (MPI_Comm, PetscInt m, PetscInt n, PetscInt M, PetscInt N,
              PetscInt, const PetscInt[], PetscInt, const PetscInt[], Mat *A)
{
  const PetscInt rows = (M >= 0) ? M : m, cols = (N >= 0) ? N : n;
  PetscCheck (rows >= 0 && cols >= 0, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "MatCreateAIJ: sizes");
  *A = mat_new (LGH_MAT_AIJ, rows, cols);
  (*A)->rows.resize ((size_t) rows);
  return 0;
}

PetscErrorCode
MatCreateSeqAIJ (MPI_Comm c, PetscInt m, PetscInt n, PetscInt nz,
                 const PetscInt nnz[], Mat *A)
{
  return MatCreateAIJ (c, m, n, m, n, nz, nnz, 0, NULL, A);
}

PetscErrorCode
MatCreateSeqAIJWithArrays (MPI_Comm, PetscInt m, PetscInt n, PetscInt i[],
                           PetscInt j[], PetscScalar a[], Mat *A)
{
  Mat M = mat_new (LGH_MAT_AIJ, m, n);
  M->rp.assign (i, i + m + 1);
  M->ci.assign (j, j + i[m]);
  M->v.assign (a, a + i[m]);
  M->assembled = 1;
  *A = M;
  return 0;
}

PetscErrorCode
MatCreateDense (MPI_Comm, PetscInt m, PetscInt n, PetscInt M, PetscInt N,
                PetscScalar *data, Mat *A)
{
  const PetscInt rows = (M >= 0) ? M : m, cols = (N >= 0) ? N : n;
  Mat D = mat_new (LGH_MAT_DENSE, rows, cols);
  D->lda = rows;
  if (data) D->a = data;
  else { D->dense.assign ((size_t) rows * cols, 0.0); D->a = D->dense.data (); }
  D->assembled = 1;
  *A = D;
  return 0;
}

PetscErrorCode
MatCreateSeqDense (MPI_Comm c, PetscInt m, PetscInt n, PetscScalar *data,
                   Mat *A)
{
  return MatCreateDense (c, m, n, m, n, data, A);
}

PetscErrorCode
MatSetValue (Mat A, PetscInt i, PetscInt j, PetscScalar v, InsertMode mode)
{
  PetscCheck (i >= 0 && i < A->m && j >= 0 && j < A->n, 0,
              PETSC_ERR_ARG_OUTOFRANGE, "MatSetValue: (%d,%d) out of range",
              (int) i, (int) j);
  A->state++;
  if (A->type == LGH_MAT_DENSE) {
    double &x = A->a[i + (size_t) j * A->lda];
    x = (mode == ADD_VALUES) ? x + v : v;
    return 0;
  }
  if (A->assembled && A->rows.empty ()) {
    // reopen an assembled AIJ: move CSR back into the row builder
    A->rows.resize ((size_t) A->m);
    for (int r = 0; r < A->m; r++)
      for (int k = A->rp[r]; k < A->rp[r + 1]; k++) A->rows[r][A->ci[k]] = A->v[k];
  }
  A->assembled = 0;
  auto &row = A->rows[i];
  auto it = row.find (j);
  if (it == row.end ()) row[j] = v;
  else it->second = (mode == ADD_VALUES) ? it->second + v : v;
  return 0;
}

PetscErrorCode
MatSetValues (Mat A, PetscInt m, const PetscInt idxm[], PetscInt n,
              const PetscInt idxn[], const PetscScalar v[], InsertMode mode)
{
  for (int r = 0; r < m; r++)
    for (int c = 0; c < n; c++)
      PetscCall (MatSetValue (A, idxm[r], idxn[c], v[(size_t) r * n + c], mode));
  return 0;
}

PetscErrorCode MatAssemblyBegin (Mat, MatAssemblyType) { return 0; }

PetscErrorCode
MatAssemblyEnd (Mat A, MatAssemblyType)
{
  if (A->type != LGH_MAT_AIJ || A->assembled) return 0;
  A->rp.assign ((size_t) A->m + 1, 0);
  A->ci.clear ();
  A->v.clear ();
  for (int r = 0; r < A->m; r++) {
    for (auto &kv : A->rows[r]) { A->ci.push_back (kv.first); A->v.push_back (kv.second); }
    A->rp[r + 1] = (int) A->ci.size ();
  }
  A->rows.clear ();
  A->rows.shrink_to_fit ();
  A->assembled = 1;
  A->state++;
  return 0;
}

PetscErrorCode MatGetSize (Mat A, PetscInt *M, PetscInt *N)
{ if (M) *M = A->m; if (N) *N = A->n; return 0; }
PetscErrorCode MatGetLocalSize (Mat A, PetscInt *m, PetscInt *n)
{ if (m) *m = A->m; if (n) *n = A->n; return 0; }
PetscErrorCode MatGetOwnershipRange (Mat A, PetscInt *lo, PetscInt *hi)
{ if (lo) *lo = 0; if (hi) *hi = A->m; return 0; }
PetscErrorCode MatGetType (Mat A, MatType *t)
{ *t = (A->type == LGH_MAT_AIJ) ? MATAIJ : MATDENSE; return 0; }

PetscErrorCode
MatMult (Mat A, Vec x, Vec y)
{
  PetscCheck (x->n == A->n && y->n == A->m, 0, PETSC_ERR_ARG_SIZ,
              "MatMult: size mismatch");
  PetscCheck (A->assembled, 0, PETSC_ERR_ARG_WRONGSTATE,
              "MatMult: matrix not assembled");
  if (A->type == LGH_MAT_AIJ) {
    for (int i = 0; i < A->m; i++) {
      double s = 0.0;
      for (int k = A->rp[i]; k < A->rp[i + 1]; k++) s += A->v[k] * x->a[A->ci[k]];
      y->a[i] = s;
    }
  }
  else {
    for (int i = 0; i < A->m; i++) y->a[i] = 0.0;
    for (int j = 0; j < A->n; j++) {
      const double xj = x->a[j];
      const double *col = A->a + (size_t) j * A->lda;
      for (int i = 0; i < A->m; i++) y->a[i] += col[i] * xj;
    }
  }
  return 0;
}

PetscErrorCode
MatGetDiagonal (Mat A, Vec d)
{
  for (int i = 0; i < A->m; i++) {
    double s = 0.0;
    if (A->type == LGH_MAT_AIJ) {
      for (int k = A->rp[i]; k < A->rp[i + 1]; k++) if (A->ci[k] == i) s += A->v[k];
    }
    else s = A->a[i + (size_t) i * A->lda];
    d->a[i] = s;
  }
  return 0;
}

PetscErrorCode
MatCreateVecs (Mat A, Vec *right, Vec *left)
{
  if (right) PetscCall (vec_new (A->n, right));
  if (left) PetscCall (vec_new (A->m, left));
  return 0;
}

PetscErrorCode
MatDestroy (Mat *A)
{
  if (!A || !*A) return 0;
  if (--(*A)->refct <= 0) delete *A;
  *A = NULL;
  return 0;
}

static PetscErrorCode
need_dense (Mat A)
{
  PetscCheck (A->type == LGH_MAT_DENSE, 0, PETSC_ERR_ARG_WRONG,
              "MatDense*: not a dense matrix");
  return 0;
}

PetscErrorCode MatDenseGetLDA (Mat A, PetscInt *lda)
{ PetscCall (need_dense (A)); *lda = A->lda; return 0; }
PetscErrorCode MatDenseGetArray (Mat A, PetscScalar **a)
{ PetscCall (need_dense (A)); A->state++; *a = A->a; return 0; }
PetscErrorCode MatDenseRestoreArray (Mat A, PetscScalar **a)
{ A->state++; *a = NULL; return 0; }
PetscErrorCode MatDenseGetArrayRead (Mat A, const PetscScalar **a)
{ PetscCall (need_dense (A)); *a = A->a; return 0; }
PetscErrorCode MatDenseRestoreArrayRead (Mat, const PetscScalar **a)
{ *a = NULL; return 0; }
PetscErrorCode MatDenseGetArrayWrite (Mat A, PetscScalar **a)
{ PetscCall (need_dense (A)); A->state++; *a = A->a; return 0; }
PetscErrorCode MatDenseRestoreArrayWrite (Mat A, PetscScalar **a)
{ A->state++; *a = NULL; return 0; }

PetscErrorCode
MatDenseGetColumnVecRead (Mat A, PetscInt j, Vec *v)
{
  PetscCall (need_dense (A));
  return VecCreateSeqWithArray (0, 1, A->m, A->a + (size_t) j * A->lda, v);
}

PetscErrorCode MatDenseRestoreColumnVecRead (Mat, PetscInt, Vec *v)
{ return VecDestroy (v); }

PetscErrorCode
MatDenseGetColumnVecWrite (Mat A, PetscInt j, Vec *v)
{
  A->state++;
  return MatDenseGetColumnVecRead (A, j, v);
}

PetscErrorCode MatDenseRestoreColumnVecWrite (Mat A, PetscInt, Vec *v)
{ A->state++; return VecDestroy (v); }

PetscErrorCode
MatSeqAIJGetCSR (Mat A, const PetscInt **i, const PetscInt **j,
                 const PetscScalar **a)
{
  PetscCheck (A->type == LGH_MAT_AIJ && A->assembled, 0,
              PETSC_ERR_ARG_WRONGSTATE, "MatSeqAIJGetCSR: assembled AIJ only");
  *i = A->rp.data (); *j = A->ci.data (); *a = A->v.data ();
  return 0;
}

PetscErrorCode KSPDestroy (KSP *ksp) { if (ksp) *ksp = NULL; return 0; }

}  // extern "C"

// ---- device mirror -----------------------------------------------------------
_p_Mat::~_p_Mat ()
{
  delete dev[0];
  delete dev[1];
}

namespace lgh {

int
mat_host_csr (Mat A, const int **rp, const int **ci, const double **v,
              std::vector<int> &rps, std::vector<int> &cis,
              std::vector<double> &vs)
{
  LGH_CHECK (A->assembled, 9, "matrix not assembled");
  if (A->type == LGH_MAT_AIJ) {
    *rp = A->rp.data (); *ci = A->ci.data (); *v = A->v.data ();
    return 0;
  }
  // dense: every stored entry becomes a CSR entry (zeros included)
  rps.assign ((size_t) A->m + 1, 0);
  cis.resize ((size_t) A->m * A->n);
  vs.resize ((size_t) A->m * A->n);
  for (int i = 0; i < A->m; i++) {
    for (int j = 0; j < A->n; j++) {
      cis[(size_t) i * A->n + j] = j;
      vs[(size_t) i * A->n + j] = A->a[i + (size_t) j * A->lda];
    }
    rps[i + 1] = (i + 1) * A->n;
  }
  *rp = rps.data (); *ci = cis.data (); *v = vs.data ();
  return 0;
}

int
mat_device_csr (Mat A, int prec, int want_dinv, DevCSR **out)
{
  if (A->dev[prec] && A->dev_state[prec] == A->state
      && (!want_dinv || A->dev[prec]->dinv)) {
    *out = A->dev[prec];
    return 0;
  }
  delete A->dev[prec];
  A->dev[prec] = nullptr;
  const int *rp, *ci;
  const double *v;
  std::vector<int> rps, cis;
  std::vector<double> vs;
  LGH_TRY (mat_host_csr (A, &rp, &ci, &v, rps, cis, vs));
  // values that are all exact in FP32 (an operator assembled in single
  // precision) are stored in FP32 for FP64 sweeps too: the SpMM widens them
  // exactly, so the product is bit-identical and reads 4 fewer bytes per
  // entry.  Matrices that need the inverse diagonal (Z, the smoothed
  // operators) keep their precision.
  int vprec = prec;
  if (prec == F64 && !want_dinv) {
    const int64_t nnz = rp[A->m];
    bool exact = nnz > 0;
    for (int64_t k = 0; k < nnz && exact; k++)
      exact = (double) (float) v[k] == v[k];
    if (exact) vprec = F32;
  }
  DevCSR *d = nullptr;
  LGH_TRY (csr_upload (A->m, A->n, rp, ci, v, vprec, want_dinv, &d));
  // wide rows (B: a PSF row has tens to hundreds of entries, mostly shared
  // with the neighbouring rows) also get the row-grouped copy its SpMM
  // reads, when the matrix is large enough for the product to matter
  // (building the copy costs host time of the order of one small build)
  // or a configuration is forced (tests of every candidate on small
  // matrices); stencil operators (Z, the AMG levels) do not
  if (!want_dinv && A->m > 0 && rp[A->m] >= 16 * (int64_t) A->m
      && (rp[A->m] >= ((int64_t) 1 << 24) || spmm_group_forced ())) {
    const int rc = csr_group_rows (*d, rp, ci, v);
    if (rc) { delete d; return rc; }
  }
  A->dev[prec] = d;
  A->dev_state[prec] = A->state;
  *out = d;
  return 0;
}

}  // namespace lgh
