/* petscksp.h — the KSP type is opaque in the PETSc subset: the engine has
 * no KSP implementation (lgh_prior_create_ksp reports PETSC_ERR_SUP). */
#ifndef LGH_PETSCKSP_H
#define LGH_PETSCKSP_H
#include "petscmat.h"
#ifdef __cplusplus
extern "C" {
#endif
typedef struct _p_KSP *KSP;
typedef struct _p_PC *PC;
PetscErrorCode KSPDestroy (KSP *ksp);
#ifdef __cplusplus
}
#endif
#endif
