/* PETSc is not used: this header gives reference-style programs the C library
 * headers that petsc.h brings in, and the QuaC declarations. */
#ifndef QUAC_COMPAT_PETSC_H
#define QUAC_COMPAT_PETSC_H
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "quac.h"
#ifndef PETSC_PI
#define PETSC_PI 3.14159265358979323846264338327950288
#endif
/* |x| of a PetscScalar (double complex) */
#define PetscAbsScalar(x) cabs(x)
#endif
