/* solver.c -- steady state of a qsystem (solver.c:33-220 semantics).
 *
 * The reference offers steady_state(Vec) only for its legacy global-operator
 * API (provided over a process-wide qsystem in legacy.c); here it works on a
 * qsystem: the trace row is added to row 0 of the
 * Liouvillian, b = e0, GMRES(restart) with rtol 1e-11 by default
 * (solver.c:13-15), on the device in FP64.  The initial guess is whatever the
 * qvec holds (zero for a fresh qvec, matching the reference's effective x=0).
 * The reference preconditions with PCASM/ILU(0) (solver.c:182-183); here the
 * default is the Hamiltonian-eigenbasis operator of qh_precond.hip (Jacobi and
 * none are selectable). */
#include <math.h>

#include "quac_internal.h"

static void solve(qsystem qsys, qvec x, qvec u, PetscInt restart, PetscReal tol, PetscInt maxit, PetscInt precond,
                  PetscInt *iters, PetscReal *residual) {
  if (!qsys->mat_allocated) QUAC_FAIL(0, "Matrix does not seem to be constructed! You need to call construct_matrix!");
  if (!qsys->dm_equations || qsys->mcwf_solver)
    QUAC_FAIL(0, "Steady state is only supported for density-matrix (Lindblad) systems");
  if (x->my_type != DENSITY_MATRIX || x->n != qsys->dim) QUAC_FAIL(0, "steady_state_sys needs a density matrix");
  if (precond < QUAC_SS_PC_NONE || precond > QUAC_SS_PC_HBASIS) QUAC_FAIL(0, "unknown steady-state preconditioner");
  if (u && (u->my_type != DENSITY_MATRIX || u->n != qsys->dim))
    QUAC_FAIL(0, "steady_state_sys_aug: the predicted state must be a density matrix of the system");
  Vec uv = u ? u->data : NULL;
  int64_t it = 0;
  double res = 0.0;
  int rc = 0;
  if (precond == QUAC_SS_PC_HBASIS) {
    /* the eigenbasis of the effective Hamiltonian's Hermitian part (rebuilt per
     * solve: the terms may have changed since the last one) */
    Mat heff = NULL;
    qh_pc *pc = NULL;
    if (quac_assemble_heff(qsys, &heff)) return;
    rc = qh_pc_create_hbasis(heff, &pc);
    qh_mat_destroy(heff);
    if (!rc) rc = qh_gmres_aug(qsys->mat_A, qsys->total_levels, x->data, uv, (int)restart, tol, maxit, QH_PC_HBASIS,
                               pc, &it, &res);
    qh_pc_destroy(pc);
  } else {
    rc = qh_gmres_aug(qsys->mat_A, qsys->total_levels, x->data, uv, (int)restart, tol, maxit, (int)precond, NULL, &it,
                      &res);
  }
  if (rc) QUAC_FAIL(9, "steady state solve failed: %s", qh_last_error());
  if (iters) *iters = it;
  if (residual) *residual = res;
  if (quac_opt_int("quac_verbose", 0)) PetscPrintf(0, "Iterations %lld residual %e\n", (long long)it, res);
}

void steady_state_sys_opts(qsystem qsys, qvec x, PetscInt restart, PetscReal tol, PetscInt maxit, PetscInt precond,
                           PetscInt *iters, PetscReal *residual) {
  solve(qsys, x, NULL, restart, tol, maxit, precond, iters, residual);
}

/* the solve with its search space augmented by a predicted steady state u
 * (qh_gmres_aug); x holds the initial guess as in steady_state_sys_opts */
void steady_state_sys_aug(qsystem qsys, qvec x, qvec u, PetscInt restart, PetscReal tol, PetscInt maxit,
                          PetscInt precond, PetscInt *iters, PetscReal *residual) {
  solve(qsys, x, u, restart, tol, maxit, precond, iters, residual);
}

/* defaults: restart 100, rtol 1e-11, at most 10000 iterations (solver.c:13-15,
 * 186-187); the preconditioner is the Hamiltonian-eigenbasis operator unless
 * -quac_ss_pc selects 0 (none) or 1 (Jacobi) */
void steady_state_sys(qsystem qsys, qvec x) {
  PetscInt restart = quac_opt_int("ksp_gmres_restart", 100);
  PetscReal tol = quac_opt_double("ksp_rtol", 1e-11);
  PetscInt maxit = quac_opt_int("ksp_max_it", 10000);
  PetscInt pc = quac_opt_int("quac_ss_pc", QUAC_SS_PC_HBASIS);
  steady_state_sys_opts(qsys, x, restart, tol, maxit, pc, NULL, NULL);
}
