/* solver.c -- steady state of a qsystem (solver.c:33-220 semantics).
 *
 * The reference offers steady_state(Vec) only for its legacy global-operator
 * API; here it works on a qsystem: the trace row is added to row 0 of the
 * Liouvillian, b = e0, GMRES(restart) with rtol 1e-11 by default
 * (solver.c:13-15), on the device in FP64.  The initial guess is whatever the
 * qvec holds (zero for a fresh qvec, matching the reference's effective x=0). */
#include <math.h>

#include "quac_internal.h"

void steady_state_sys_opts(qsystem qsys, qvec x, PetscInt restart, PetscReal tol, PetscInt maxit, PetscBool jacobi,
                           PetscInt *iters, PetscReal *residual) {
  if (!qsys->mat_allocated) QUAC_FAIL(0, "Matrix does not seem to be constructed! You need to call construct_matrix!");
  if (!qsys->dm_equations || qsys->mcwf_solver)
    QUAC_FAIL(0, "Steady state is only supported for density-matrix (Lindblad) systems");
  if (x->my_type != DENSITY_MATRIX || x->n != qsys->dim) QUAC_FAIL(0, "steady_state_sys needs a density matrix");
  int64_t it = 0;
  double res = 0.0;
  int rc = qh_gmres(qsys->mat_A, qsys->total_levels, x->data, (int)restart, tol, maxit, jacobi ? 1 : 0, &it, &res);
  if (rc) QUAC_FAIL(9, "steady state solve failed: %s", qh_last_error());
  if (iters) *iters = it;
  if (residual) *residual = res;
  if (quac_opt_int("quac_verbose", 0)) PetscPrintf(0, "Iterations %lld residual %e\n", (long long)it, res);
}

void steady_state_sys(qsystem qsys, qvec x) {
  PetscInt restart = quac_opt_int("ksp_gmres_restart", 100);
  PetscReal tol = quac_opt_double("ksp_rtol", 1e-11);
  PetscInt maxit = quac_opt_int("ksp_max_it", 10000);
  PetscBool jac = quac_opt_int("quac_jacobi", 1) ? PETSC_TRUE : PETSC_FALSE;
  steady_state_sys_opts(qsys, x, restart, tol, maxit, jac, NULL, NULL);
}
