/* petscvec.h — sequential Vec of the PETSc subset (see petscsys.h). */
#ifndef LGH_PETSCVEC_H
#define LGH_PETSCVEC_H
#include "petscsys.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct _p_Vec *Vec;
typedef enum { NORM_1 = 0, NORM_2 = 1, NORM_INFINITY = 3 } NormType;

PetscErrorCode VecCreateSeq (MPI_Comm comm, PetscInt n, Vec *v);
PetscErrorCode VecCreateSeqWithArray (MPI_Comm comm, PetscInt bs, PetscInt n,
                                      const PetscScalar *array, Vec *v);
PetscErrorCode VecCreateMPI (MPI_Comm comm, PetscInt n, PetscInt N, Vec *v);
PetscErrorCode VecDuplicate (Vec x, Vec *y);
PetscErrorCode VecDestroy (Vec *v);
PetscErrorCode VecCopy (Vec x, Vec y);
PetscErrorCode VecSet (Vec x, PetscScalar a);
PetscErrorCode VecZeroEntries (Vec x);
PetscErrorCode VecScale (Vec x, PetscScalar a);
PetscErrorCode VecAXPY (Vec y, PetscScalar a, Vec x);
PetscErrorCode VecNorm (Vec x, NormType t, PetscReal *nrm);
PetscErrorCode VecDot (Vec x, Vec y, PetscScalar *d);
PetscErrorCode VecPointwiseMult (Vec w, Vec x, Vec y);
PetscErrorCode VecPointwiseDivide (Vec w, Vec x, Vec y);
PetscErrorCode VecSqrtAbs (Vec x);
PetscErrorCode VecReciprocal (Vec x);
PetscErrorCode VecGetSize (Vec x, PetscInt *n);
PetscErrorCode VecGetLocalSize (Vec x, PetscInt *n);
PetscErrorCode VecGetOwnershipRange (Vec x, PetscInt *lo, PetscInt *hi);
PetscErrorCode VecGetArray (Vec x, PetscScalar **a);
PetscErrorCode VecRestoreArray (Vec x, PetscScalar **a);
PetscErrorCode VecGetArrayRead (Vec x, const PetscScalar **a);
PetscErrorCode VecRestoreArrayRead (Vec x, const PetscScalar **a);
PetscErrorCode VecGetArrayWrite (Vec x, PetscScalar **a);
PetscErrorCode VecRestoreArrayWrite (Vec x, PetscScalar **a);

#ifdef __cplusplus
}
#endif
#endif
