#ifndef LGH_PETSC_H
#define LGH_PETSC_H
#include "petscsys.h"
#include "petscvec.h"
#include "petscmat.h"
#include "petscksp.h"
#endif
