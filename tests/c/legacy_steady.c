/*
 * Legacy steady_state (solver.c:33-221) of a driven, detuned, damped qubit:
 * H = omega/2 sigma_x + delta n, L = gamma L[sigma]; prints rho_ee, whose
 * closed form is (omega^2/4) / (delta^2 + gamma^2/4 + omega^2/2), and the trace.
 */
#include <stdio.h>

#include "quac_legacy.h"

int main(int argc, char **args) {
  operator q;
  Vec rho;
  PetscScalar ee, tr;
  QuaC_initialize(argc, args);
  create_op(2, &q);
  add_to_ham(0.2, q->sig_x);
  add_to_ham(0.3, q->n);
  add_lin(1.0, q);
  create_full_dm(&rho);
  steady_state(rho);
  get_dm_element(rho, 1, 1, &ee);
  trace_dm(&tr, rho);
  printf("steady %.15e %.15e\n", PetscRealPart(ee), PetscRealPart(tr));
  destroy_dm(rho);
  QuaC_finalize();
  return 0;
}
