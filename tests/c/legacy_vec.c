/*
 * Legacy API with a VEC (three-level) subsystem: |1> decays to |0> through the
 * Lindblad operator |0><1| (add_lin_mult2 of two VEC operators,
 * operators.c:743-821) while a cavity is coupled only through a VEC-pair
 * Hamiltonian term that never acts (|2><2| energy, add_to_ham_mult2).  The
 * initial populations (1/4 in |2>, 3/4 in |1>) are given per VEC level
 * (set_initial_pop, dm_utilities.c:665-804).  Prints the four populations
 * (three VEC levels, cavity <n>) at the end.
 */
#include <stdio.h>
#include <stdlib.h>

#include "quac_legacy.h"

int main(int argc, char **args) {
  vec_op atom;
  operator cav;
  Vec rho;
  double *pops;
  QuaC_initialize(argc, args);
  create_vec(3, &atom);
  create_op(3, &cav);
  add_to_ham_mult2(0.7, atom[2], atom[2]);
  add_to_ham(0.3, cav->n);
  add_lin_mult2(0.5, atom[0], atom[1]);
  set_initial_pop(atom[1], 3.0);
  set_initial_pop(atom[2], 1.0);
  pops = malloc(sizeof(double) * get_num_populations());
  create_full_dm(&rho);
  set_dm_from_initial_pop(rho);
  time_step(rho, 0.0, 2.0, 0.01, 100000);
  get_populations(rho, &pops);
  printf("pops %d %.15e %.15e %.15e %.15e\n", get_num_populations(), pops[0], pops[1], pops[2], pops[3]);
  destroy_dm(rho);
  free(pops);
  QuaC_finalize();
  return 0;
}
