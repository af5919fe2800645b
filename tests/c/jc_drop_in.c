/*
 * Drop-in check: a QuaC user program written against the reference API
 * (the call sequence of examples/simple_jc_test.c:19-91, with the config of a
 * Jaynes-Cummings emitter + 10-level cavity), compiled against include/quac.h
 * and linked with libquac.so instead of PETSc-based QuaC.
 *
 * Prints "step t <sigma^dag sigma> <a^dag a>" every 100 steps.
 */
#include <math.h>
#include <stdio.h>

#include "quac.h"

static operator sigma, b;
static qvec dm_dummy;

PetscErrorCode ts_monitor(TS ts, PetscInt step, PetscReal time, Vec rho, void *ctx) {
  Vec tmp_data;
  PetscScalar e1, e2;
  (void)ts;
  (void)ctx;
  if (step % 100 != 0) PetscFunctionReturn(0);
  /* the reference's aliasing idiom: put the stepper's Vec into a dummy qvec */
  tmp_data = dm_dummy->data;
  dm_dummy->data = rho;
  get_expectation_value_qvec(dm_dummy, &e1, 1, sigma->n);
  get_expectation_value_qvec(dm_dummy, &e2, 1, b->n);
  PetscPrintf(PETSC_COMM_WORLD, "%lld %.12f %.15e %.15e\n", (long long)step, time, PetscRealPart(e1),
              PetscRealPart(e2));
  dm_dummy->data = tmp_data;
  PetscFunctionReturn(0);
}

int main(int argc, char **args) {
  qvec rho;
  qsystem qsys;
  PetscScalar tr;
  PetscReal wq = 1.0, wc = 1.0, g = 0.1, kappa = 1e-2, gamma = 1e-2;
  PetscReal time_max = 10.0, dt = 0.01;
  PetscInt steps_max = 100000;

  QuaC_initialize(argc, args);
  initialize_system(&qsys);
  create_op_sys(qsys, 2, &sigma);
  create_op_sys(qsys, 10, &b);

  add_ham_term(qsys, wq, 1, sigma->n);
  add_ham_term(qsys, wc, 1, b->n);
  add_ham_term(qsys, g, 2, sigma->dag, b);
  add_ham_term(qsys, g, 2, b->dag, sigma);
  add_lin_term(qsys, gamma, 1, sigma);
  add_lin_term(qsys, kappa, 1, b);
  construct_matrix(qsys);

  set_ts_monitor_sys(qsys, ts_monitor, NULL);
  create_qvec_sys(qsys, &rho);
  create_qvec_sys(qsys, &dm_dummy);
  /* emitter excited, cavity vacuum: Hilbert index 1*10 + 0 */
  add_to_qvec(rho, 1.0, (PetscInt)10, (PetscInt)10);
  assemble_qvec(rho);

  time_step_sys(qsys, rho, 0.0, time_max, dt, steps_max);
  get_trace_qvec(rho, &tr);
  PetscPrintf(PETSC_COMM_WORLD, "trace %.15e\n", PetscRealPart(tr));

  destroy_qvec(&rho);
  destroy_qvec(&dm_dummy);
  destroy_op_sys(&sigma);
  destroy_op_sys(&b);
  destroy_system(&qsys);
  QuaC_finalize();
  return 0;
}
