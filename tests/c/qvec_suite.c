/*
 * The state-container and observable cases of the reference's
 * tests/qvec_utilities_tests.c, written against include/quac.h with the same
 * calls, struct field reads and expected values:
 *   creation and dimensions (:13-230), expectation values of a wavefunction and
 *   a density matrix (:232-324), the random wavefunction fixture (:872-899, path
 *   in argv[1]), bitstring probabilities of a wavefunction, a density matrix and
 *   an MCWF ensemble (:901-1020).
 * Prints "FAIL <case> <detail>" per failed check and "checks <n> failures <m>".
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "quac.h"

static int n_checks = 0, n_fail = 0;

static void check(int ok, const char *name, const char *what, double got, double want) {
  n_checks++;
  if (!ok) {
    n_fail++;
    printf("FAIL %s %s got %.17g want %.17g\n", name, what, got, want);
  }
}
#define EQ_INT(name, got, want) check((long long)(got) == (long long)(want), name, #got, (double)(got), (double)(want))
/* the reference's TEST_ASSERT_EQUAL_FLOAT: relative 1e-5, here 1e-12 absolute for these small values */
#define EQ_FLT(name, got, want) check(fabs((double)(got) - (double)(want)) <= 1e-12, name, #got, (double)(got), (double)(want))

static void creation(void) {
  qvec wf, dm;
  create_arb_qvec(&wf, 144, WAVEFUNCTION);
  create_arb_qvec(&dm, 144, DENSITY_MATRIX);
  EQ_INT("create_arb_qvec", wf->n, 144);
  EQ_INT("create_arb_qvec", wf->total_levels, 144);
  EQ_INT("create_arb_qvec", dm->n, 144);
  EQ_INT("create_arb_qvec", dm->total_levels, 12);
  destroy_qvec(&wf);
  destroy_qvec(&dm);

  PetscInt dims[4] = {2, 3, 4, 2}, dims_dm[8] = {2, 3, 4, 2, 2, 3, 4, 2};
  create_arb_qvec_dims(&wf, 4, dims, WAVEFUNCTION);
  create_arb_qvec_dims(&dm, 8, dims_dm, DENSITY_MATRIX);
  EQ_INT("create_arb_qvec_dims", wf->n, 48);
  EQ_INT("create_arb_qvec_dims", wf->total_levels, 48);
  EQ_INT("create_arb_qvec_dims", dm->n, 48 * 48);
  EQ_INT("create_arb_qvec_dims", dm->total_levels, 48);
  destroy_qvec(&wf);
  destroy_qvec(&dm);

  create_arb_qvec(&wf, 48, WAVEFUNCTION);
  change_qvec_dims(wf, 4, dims);
  EQ_INT("change_qvec_dims", wf->n_ops, 4);
  EQ_INT("change_qvec_dims", wf->ndims_hspace, 4);
  for (int i = 0; i < 4; i++) EQ_INT("change_qvec_dims", wf->hspace_dims[i], dims[i]);
  create_arb_qvec(&dm, 48 * 48, DENSITY_MATRIX);
  change_qvec_dims(dm, 8, dims_dm);
  EQ_INT("change_qvec_dims", dm->n_ops, 4);
  EQ_INT("change_qvec_dims", dm->ndims_hspace, 8);
  for (int i = 0; i < 8; i++) EQ_INT("change_qvec_dims", dm->hspace_dims[i], dims_dm[i]);
  destroy_qvec(&wf);
  destroy_qvec(&dm);

  /* qvecs of a system of levels 2, 3, 4 (n = 24) */
  for (int kind = 0; kind < 5; kind++) {
    qsystem qsys;
    operator op2, op3, op4;
    qvec q;
    initialize_system(&qsys);
    create_op_sys(qsys, 2, &op2);
    create_op_sys(qsys, 3, &op3);
    create_op_sys(qsys, 4, &op4);
    if (kind == 1) {
      add_lin_term(qsys, 1.0, 1, op3);
      use_mcwf_solver(qsys, 10, 0);
    }
    if (kind == 2) add_lin_term(qsys, 0.0, 1, op3);
    if (kind == 3) create_wf_sys(qsys, &q);
    else if (kind == 4) create_dm_sys(qsys, &q);
    else create_qvec_sys(qsys, &q);
    const int dmk = kind == 2 || kind == 4;
    const char *name = kind == 0 ? "create_qvec_sys_wf" : kind == 1 ? "create_qvec_sys_wf_ens" :
                       kind == 2 ? "create_qvec_sys_dm" : kind == 3 ? "create_wf_sys" : "create_dm_sys";
    EQ_INT(name, q->n, dmk ? 24 * 24 : 24);
    EQ_INT(name, q->total_levels, 24);
    EQ_INT(name, q->my_type, kind == 1 ? WF_ENSEMBLE : dmk ? DENSITY_MATRIX : WAVEFUNCTION);
    if (kind == 1) EQ_INT(name, q->n_ensemble, 10);
    destroy_op_sys(&op2);
    destroy_op_sys(&op3);
    destroy_op_sys(&op4);
    destroy_qvec(&q);
    destroy_system(&qsys);
  }
}

static void expectation(int dm) {
  const char *name = dm ? "get_expectation_value_dm_real" : "get_expectation_value_wf_real";
  operator qd1, qd2;
  PetscScalar ev1, ev2, ev3, ev4;
  qvec q;
  qsystem qsys;
  initialize_system(&qsys);
  create_op_sys(qsys, 2, &qd1);
  create_op_sys(qsys, 2, &qd2);
  if (dm) create_dm_sys(qsys, &q);
  else create_wf_sys(qsys, &q);
  add_to_qvec_fock_op(1.0, q, 2, qd1, (PetscInt)1, qd2, (PetscInt)1);
  assemble_qvec(q);
  get_expectation_value_qvec(q, &ev1, 1, qd1->n);
  get_expectation_value_qvec(q, &ev2, 1, qd2->n);
  get_expectation_value_qvec(q, &ev3, 2, qd1->dag, qd2);
  get_expectation_value_qvec(q, &ev4, 2, qd2->dag, qd1);
  EQ_FLT(name, PetscRealPart(ev1), 1.0);
  EQ_FLT(name, PetscImaginaryPart(ev1), 0.0);
  EQ_FLT(name, PetscRealPart(ev2), 1.0);
  EQ_FLT(name, PetscImaginaryPart(ev2), 0.0);
  EQ_FLT(name, PetscRealPart(ev3), 0.0);
  EQ_FLT(name, PetscImaginaryPart(ev3), 0.0);
  EQ_FLT(name, PetscRealPart(ev4), 0.0);
  EQ_FLT(name, PetscImaginaryPart(ev4), 0.0);
  destroy_op_sys(&qd1);
  destroy_op_sys(&qd2);
  destroy_system(&qsys);
  destroy_qvec(&q);
}

static void read_wf(const char *path) {
  static const double want[4][2] = {{-0.637991927277723, 0.3126100254765921},
                                    {0.08900411349251001, -0.19371936261879386},
                                    {0.45877685279777203, 0.06902325924823785},
                                    {0.1497653947732421, 0.4605673290368073}};
  qvec wf;
  PetscScalar val;
  read_qvec_wf_binary(&wf, path);
  for (int i = 0; i < 4; i++) {
    get_wf_element_qvec(wf, i, &val);
    EQ_FLT("read_wf_binary", PetscRealPart(val), want[i][0]);
    EQ_FLT("read_wf_binary", PetscImaginaryPart(val), want[i][1]);
  }
  destroy_qvec(&wf);
}

static void bitstrings(int kind) {
  const char *name = kind == 0 ? "get_bitstring_probs_wf" : kind == 1 ? "get_bitstring_probs_dm" :
                     "get_bitstring_probs_wf_ens";
  qvec q;
  qsystem qsys;
  operator op0, op1, op2;
  PetscInt nloc;
  PetscReal *probs, *vars;
  initialize_system(&qsys);
  create_op_sys(qsys, 2, &op0);
  create_op_sys(qsys, 2, &op1);
  create_op_sys(qsys, 2, &op2);
  if (kind >= 1) add_lin_term(qsys, kind == 1 ? 1.0 / 8.0 : 1.0, 1, op0);
  if (kind == 2) use_mcwf_solver(qsys, 8, 0);
  create_qvec_sys(qsys, &q);
  for (PetscInt i = 0; i < 8; i++) {
    if (kind == 0) add_to_qvec(q, 1 / sqrt(8), i);
    else if (kind == 1) add_to_qvec(q, 1.0 / 8.0, i, i);
    else add_to_wf_ens_loc(q, i, 1.0, i);
  }
  assemble_qvec(q);
  get_bitstring_probs(q, &nloc, &probs, &vars);
  EQ_INT(name, nloc, 8);
  for (int i = 0; i < 8; i++) EQ_FLT(name, probs[i], 1.0 / 8.0);
  free(probs);
  free(vars);
  destroy_op_sys(&op0);
  destroy_op_sys(&op1);
  destroy_op_sys(&op2);
  destroy_qvec(&q);
  destroy_system(&qsys);
}

int main(int argc, char **args) {
  QuaC_initialize(1, args); /* argv[1] is the fixture path, not an option */
  creation();
  expectation(0);
  expectation(1);
  if (argc > 1) read_wf(args[1]);
  for (int k = 0; k < 3; k++) bitstrings(k);
  printf("checks %d failures %d\n", n_checks, n_fail);
  QuaC_finalize();
  return 0;
}
