/* petscsys.h — the PETSc subset the GLR API boundary needs, at
 * communicator size 1.  Type names, signatures, error codes and macro
 * semantics follow PETSc's public headers so that consumer code written
 * against the reference (tests/test_glr.cpp, tests/test_prior_mat.cpp,
 * examples/heat/heat_source_inversion.cpp) compiles unchanged.  Everything
 * is sequential host memory; the GLR engine mirrors what it needs in HBM. */
#ifndef LGH_PETSCSYS_H
#define LGH_PETSCSYS_H

#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int    PetscErrorCode;
typedef int    PetscInt;
typedef int    PetscMPIInt;
typedef double PetscScalar;
typedef double PetscReal;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef int    MPI_Comm;
typedef struct _p_PetscObject *PetscObject;

#define PETSC_SUCCESS 0
#define PETSC_COMM_WORLD ((MPI_Comm) 0x44000000)
#define PETSC_COMM_SELF  ((MPI_Comm) 0x44000001)
#define PETSC_DECIDE     (-1)
#define PETSC_DETERMINE  PETSC_DECIDE
#define PETSC_DEFAULT    (-2)
#define PETSC_NULLPTR    NULL

/* error codes (petscerror.h values) */
#define PETSC_ERR_MEM             55
#define PETSC_ERR_SUP             56
#define PETSC_ERR_ORDER           58
#define PETSC_ERR_ARG_SIZ         60
#define PETSC_ERR_ARG_OUTOFRANGE  63
#define PETSC_ERR_ARG_WRONGSTATE  73
#define PETSC_ERR_LIB             76
#define PETSC_ERR_ARG_NULL        85
#define PETSC_ERR_ARG_WRONG       62
#define PETSC_ERR_CONV_FAILED     91

PetscErrorCode PetscInitialize (int *argc, char ***argv, const char *file,
                                const char *help);
PetscErrorCode PetscFinalize (void);
PetscErrorCode PetscPrintf (MPI_Comm comm, const char *fmt, ...);
/* records the message of a failed check (retrievable through
 * lgh_petsc_last_error) and returns code                                */
PetscErrorCode lgh_petsc_error (PetscErrorCode code, const char *fmt, ...);
const char    *lgh_petsc_last_error (void);
PetscErrorCode PetscObjectReference (PetscObject obj);
PetscErrorCode PetscObjectGetComm (PetscObject obj, MPI_Comm *comm);
double         MPI_Wtime (void);

#define PetscCall(expr)                                                       \
  do { PetscErrorCode ierr_lgh_ = (PetscErrorCode) (expr);                    \
       if (ierr_lgh_) return ierr_lgh_; } while (0)
#define PetscCallMPI(expr) PetscCall (expr)
#define PetscCheck(cond, comm, code, ...)                                     \
  do { if (!(cond)) return lgh_petsc_error ((code), __VA_ARGS__); } while (0)
#define SETERRQ(comm, code, ...) return lgh_petsc_error ((code), __VA_ARGS__)
#define CHKERRABORT(comm, ierr)                                               \
  do { if (ierr) { fprintf (stderr, "PETSc-subset error %d: %s\n", (int) (ierr), \
                            lgh_petsc_last_error ()); abort (); } } while (0)
#define CHKERRQ(ierr) do { if (ierr) return ierr; } while (0)

#define PetscMax(a, b) (((a) < (b)) ? (b) : (a))
#define PetscMin(a, b) (((a) < (b)) ? (a) : (b))
#define PetscAbsReal(a) fabs (a)
#define PetscSqrtReal(a) sqrt (a)

#ifdef __cplusplus
}
#endif
#endif
