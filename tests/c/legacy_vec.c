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

#include "quac_hip.h"
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
  /* combine_ops_to_mat (operators.c:923-994): |2><2| (x) I and the identity */
  for (int k = 0; k < 2; k++) {
    Mat m = NULL;
    if (k == 0) combine_ops_to_mat(&m, total_levels, 2, atom[2], atom[2]);
    else combine_ops_to_mat(&m, total_levels, 0);
    int64_t rows = 0, nnz = 0;
    qh_mat_info(m, &rows, &nnz);
    int64_t rp[16], col[64];
    double val[128];
    qh_mat_export(m, rp, col, val);
    double tr = 0.0;
    for (int64_t r = 0; r < rows; r++)
      for (int64_t e = rp[r]; e < rp[r + 1]; e++)
        if (col[e] == r) tr += val[2 * e];
    printf("comb%d %lld %.15f\n", k, (long long)rows, tr);
    qh_mat_destroy(m);
  }
  destroy_dm(rho);
  free(pops);
  QuaC_finalize();
  return 0;
}
