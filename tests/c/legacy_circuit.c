/*
 * Legacy circuits (quantum_gates.h:68-72): a qubit decaying at rate gamma is
 * flipped by sigma_x at t = 1 (a gate added with add_gate_to_circuit to one
 * circuit, copied into a second with add_circuit_to_circuit at time 1, started
 * with start_circuit_at_time at 0), then decays: <n>(t) = 0 for t < 1 and
 * exp(-gamma (t - 1)) after.  Prints "n <n at t=2>" and "gates <count>".
 */
#include <stdio.h>

#include "quac_legacy.h"

int main(int argc, char **args) {
  operator q;
  Vec rho;
  circuit a, b;
  double pops[1];
  double *pp = pops;
  QuaC_initialize(argc, args);
  create_op(2, &q);
  add_lin(0.7, q);
  create_circuit(&a, 2);
  add_gate_to_circuit(&a, 0.25, SIGMAX, 0);
  create_circuit(&b, 2);
  add_circuit_to_circuit(&b, a, 1.0);
  start_circuit_at_time(&b, 0.0);
  create_full_dm(&rho);
  set_dm_from_initial_pop(rho);
  time_step(rho, 0.0, 2.0, 0.01, 100000);
  get_populations(rho, &pp);
  printf("n %.15e\ngates %d\n", pops[0], (int)b.num_gates);
  destroy_dm(rho);
  QuaC_finalize();
  return 0;
}
