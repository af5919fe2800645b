/*
 * A parameter sweep written against the QuaC API: eight Jaynes-Cummings systems
 * (an emitter and a cavity of 4 .. 11 levels, cavity detuned by -0.07 .. 0.07)
 * built with the reference's calls (examples/simple_jc_test.c:34-72), then
 * advanced together from t = 0 to 10 by time_step_sys_batch, and the same
 * systems advanced one at a time by time_step_sys into separate states.
 *
 * Prints "batch i <sigma^dag sigma> <a^dag a> trace" and "single i ..." per system.
 */
#include <stdio.h>

#include "quac.h"

#define NSYS 8

int main(int argc, char **args) {
  qsystem sys[NSYS];
  operator sig[NSYS], cav[NSYS];
  qvec xb[NSYS], xs[NSYS];
  QuaC_initialize(argc, args);
  for (int i = 0; i < NSYS; i++) {
    const PetscInt levels = 4 + i;
    const PetscReal wc = 1.0 + 0.02 * (i - 3.5);
    initialize_system(&sys[i]);
    create_op_sys(sys[i], 2, &sig[i]);
    create_op_sys(sys[i], levels, &cav[i]);
    add_ham_term(sys[i], 1.0, 1, sig[i]->n);
    add_ham_term(sys[i], wc, 1, cav[i]->n);
    add_ham_term(sys[i], 0.1, 2, sig[i]->dag, cav[i]);
    add_ham_term(sys[i], 0.1, 2, cav[i]->dag, sig[i]);
    add_lin_term(sys[i], 1e-2, 1, sig[i]);
    add_lin_term(sys[i], 2e-2, 1, cav[i]);
    construct_matrix(sys[i]);
    create_qvec_sys(sys[i], &xb[i]);
    create_qvec_sys(sys[i], &xs[i]);
    /* emitter excited, cavity vacuum: Hilbert index 1 * levels + 0 */
    add_to_qvec(xb[i], 1.0, levels, levels);
    add_to_qvec(xs[i], 1.0, levels, levels);
    assemble_qvec(xb[i]);
    assemble_qvec(xs[i]);
  }
  time_step_sys_batch(NSYS, sys, xb, 0.0, 10.0, 0.01, 100000);
  for (int i = 0; i < NSYS; i++) time_step_sys(sys[i], xs[i], 0.0, 10.0, 0.01, 100000);
  for (int pass = 0; pass < 2; pass++)
    for (int i = 0; i < NSYS; i++) {
      qvec x = pass ? xs[i] : xb[i];
      PetscScalar e1, e2, tr;
      get_expectation_value_qvec(x, &e1, 1, sig[i]->n);
      get_expectation_value_qvec(x, &e2, 1, cav[i]->n);
      get_trace_qvec(x, &tr);
      PetscPrintf(PETSC_COMM_WORLD, "%s %d %.15e %.15e %.15e\n", pass ? "single" : "batch", i, PetscRealPart(e1),
                  PetscRealPart(e2), PetscRealPart(tr));
    }
  for (int i = 0; i < NSYS; i++) {
    destroy_qvec(&xb[i]);
    destroy_qvec(&xs[i]);
    destroy_op_sys(&sig[i]);
    destroy_op_sys(&cav[i]);
    destroy_system(&sys[i]);
  }
  QuaC_finalize();
  return 0;
}
