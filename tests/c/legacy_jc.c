/*
 * Drop-in check of the legacy global-operator API: a Jaynes-Cummings emitter
 * + 10-level cavity written the way programs before qsystem were
 * (create_op / add_to_ham / add_to_ham_mult2 / add_lin / set_initial_pop /
 * create_full_dm / set_dm_from_initial_pop / set_ts_monitor / time_step /
 * get_populations; operators.h:43-69, dm_utilities.h:10-39, solver.h:8-15),
 * compiled against include/quac_legacy.h and linked with libquac.so.
 *
 * Prints "step t <sigma^dag sigma> <a^dag a>" every 100 steps (from
 * get_populations) and "expect <a^dag a>" and "trace" at the end.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "quac_legacy.h"

static operator qb, cav;
static double *pops;

PetscErrorCode monitor(TS ts, PetscInt step, PetscReal time, Vec rho, void *ctx) {
  (void)ts;
  (void)ctx;
  if (step % 100 != 0) return 0;
  get_populations(rho, &pops);
  printf("%lld %.12f %.15e %.15e\n", (long long)step, time, pops[0], pops[1]);
  return 0;
}

int main(int argc, char **args) {
  Vec rho;
  PetscScalar tr, ex;
  QuaC_initialize(argc, args);
  create_op(2, &qb);
  create_op(10, &cav);
  add_to_ham(1.0, qb->n);
  add_to_ham(1.0, cav->n);
  add_to_ham_mult2(0.1, qb->dag, cav);
  add_to_ham_mult2(0.1, qb, cav->dag);
  add_lin(1e-2, cav);
  add_lin(1e-2, qb);
  set_initial_pop(qb, 1);
  pops = malloc(sizeof(double) * get_num_populations());
  create_full_dm(&rho);
  set_dm_from_initial_pop(rho);
  set_ts_monitor(monitor);
  time_step(rho, 0.0, 10.0, 0.01, 100000);
  get_expectation_value(rho, &ex, 2, cav->dag, cav);
  printf("expect %.15e\n", PetscRealPart(ex));
  trace_dm(&tr, rho);
  printf("trace %.15f %d %d\n", PetscRealPart(tr), get_num_populations(), (int)total_levels);
  destroy_dm(rho);
  free(pops);
  QuaC_finalize();
  return 0;
}
