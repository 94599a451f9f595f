/* petscmat.h — sequential AIJ (CSR) and DENSE (column-major) Mat of the
 * PETSc subset (see petscsys.h). */
#ifndef LGH_PETSCMAT_H
#define LGH_PETSCMAT_H
#include "petscvec.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct _p_Mat *Mat;
typedef enum { INSERT_VALUES = 1, ADD_VALUES = 2 } InsertMode;
typedef enum { MAT_FLUSH_ASSEMBLY = 1, MAT_FINAL_ASSEMBLY = 0 } MatAssemblyType;
typedef enum { SAME_NONZERO_PATTERN = 0, DIFFERENT_NONZERO_PATTERN = 1,
               SUBSET_NONZERO_PATTERN = 2, UNKNOWN_NONZERO_PATTERN = 3 } MatStructure;
typedef const char *MatType;
#define MATAIJ   "aij"
#define MATDENSE "dense"

PetscErrorCode MatCreateAIJ (MPI_Comm comm, PetscInt m, PetscInt n, PetscInt M,
                             PetscInt N, PetscInt d_nz, const PetscInt d_nnz[],
                             PetscInt o_nz, const PetscInt o_nnz[], Mat *A);
PetscErrorCode MatCreateSeqAIJ (MPI_Comm comm, PetscInt m, PetscInt n,
                                PetscInt nz, const PetscInt nnz[], Mat *A);
PetscErrorCode MatCreateSeqAIJWithArrays (MPI_Comm comm, PetscInt m,
                                          PetscInt n, PetscInt i[],
                                          PetscInt j[], PetscScalar a[],
                                          Mat *A);
PetscErrorCode MatCreateDense (MPI_Comm comm, PetscInt m, PetscInt n,
                               PetscInt M, PetscInt N, PetscScalar *data,
                               Mat *A);
PetscErrorCode MatCreateSeqDense (MPI_Comm comm, PetscInt m, PetscInt n,
                                  PetscScalar *data, Mat *A);
PetscErrorCode MatSetValue (Mat A, PetscInt i, PetscInt j, PetscScalar v,
                            InsertMode mode);
PetscErrorCode MatSetValues (Mat A, PetscInt m, const PetscInt idxm[],
                             PetscInt n, const PetscInt idxn[],
                             const PetscScalar v[], InsertMode mode);
PetscErrorCode MatAssemblyBegin (Mat A, MatAssemblyType t);
PetscErrorCode MatAssemblyEnd (Mat A, MatAssemblyType t);
PetscErrorCode MatGetSize (Mat A, PetscInt *M, PetscInt *N);
PetscErrorCode MatGetLocalSize (Mat A, PetscInt *m, PetscInt *n);
PetscErrorCode MatGetOwnershipRange (Mat A, PetscInt *lo, PetscInt *hi);
PetscErrorCode MatMult (Mat A, Vec x, Vec y);
PetscErrorCode MatGetDiagonal (Mat A, Vec d);
PetscErrorCode MatCreateVecs (Mat A, Vec *right, Vec *left);
PetscErrorCode MatDestroy (Mat *A);
PetscErrorCode MatGetType (Mat A, MatType *t);
PetscErrorCode MatDenseGetLDA (Mat A, PetscInt *lda);
PetscErrorCode MatDenseGetArray (Mat A, PetscScalar **a);
PetscErrorCode MatDenseRestoreArray (Mat A, PetscScalar **a);
PetscErrorCode MatDenseGetArrayRead (Mat A, const PetscScalar **a);
PetscErrorCode MatDenseRestoreArrayRead (Mat A, const PetscScalar **a);
PetscErrorCode MatDenseGetArrayWrite (Mat A, PetscScalar **a);
PetscErrorCode MatDenseRestoreArrayWrite (Mat A, PetscScalar **a);
PetscErrorCode MatDenseGetColumnVecRead (Mat A, PetscInt j, Vec *v);
PetscErrorCode MatDenseRestoreColumnVecRead (Mat A, PetscInt j, Vec *v);
PetscErrorCode MatDenseGetColumnVecWrite (Mat A, PetscInt j, Vec *v);
PetscErrorCode MatDenseRestoreColumnVecWrite (Mat A, PetscInt j, Vec *v);
/* AIJ row access (CSR arrays of an assembled AIJ Mat)                   */
PetscErrorCode MatSeqAIJGetCSR (Mat A, const PetscInt **i, const PetscInt **j,
                                const PetscScalar **a);

#ifdef __cplusplus
}
#endif
#endif
