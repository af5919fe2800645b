/*
 * Legacy density-matrix utilities on two qubits a (subsystem 0) and b:
 *  - a product state rho_a (x) rho_b written with add_value_to_dm; the partial
 *    traces (partial_trace_over b, partial_trace_keep b; dm_utilities.c:150-260,
 *    468-603) and partial_trace_over_one (dm_utilities.c:1078-1117) print
 *    the 2 x 2 factors;
 *  - Werner states p |Phi+><Phi+| + (1 - p) I / 4: concurrence
 *    (dm_utilities.c:1487-1602) and fidelity to |Phi+> (dm_utilities.c:1618-1722)
 *    for p = 0.8 and p = 0.2;
 *  - the pure state cos t |01> + e^{i f} sin t |10>: concurrence sin 2t;
 *  - measure_dm (dm_utilities.c:270-308) with the projector n of a on the
 *    Werner state p = 0.8;
 *  - the sparse printers (dm_utilities.c:32-113) on |Phi+><Phi+| and on the
 *    matrix a a^dag (x) I, written to dm_sparse.txt / mat_sparse.txt in the
 *    working directory.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>

#include "quac_hip.h"
#include "quac_legacy.h"

static void put(Vec rho, const double complex *m, int n) {
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) add_value_to_dm(rho, r, c, m[n * c + r]);
}

static void show(const char *tag, Vec rho) {
  PetscScalar v;
  printf("%s", tag);
  for (int c = 0; c < 2; c++)
    for (int r = 0; r < 2; r++) {
      get_dm_element(rho, r, c, &v);
      printf(" %.15e %.15e", creal(v), cimag(v));
    }
  printf("\n");
}

static void werner(Vec rho, double p) {
  double complex m[16] = {0};
  for (int i = 0; i < 4; i++) m[5 * i] = (1.0 - p) / 4.0;
  m[0] += p / 2; /* |00><00| */
  m[15] += p / 2; /* |11><11| */
  m[12] += p / 2; /* |00><11|: row 0, column 3 */
  m[3] += p / 2;  /* |11><00| */
  put(rho, m, 4);
}

int main(int argc, char **args) {
  operator a, b;
  Vec rho, red, w, bell;
  QuaC_initialize(argc, args);
  create_op(2, &a);
  create_op(2, &b);
  add_lin(0.1, b); /* with a Lindblad term create_full_dm makes density matrices */
  const double complex ra[4] = {0.7, 0.2 - 0.1 * I, 0.2 + 0.1 * I, 0.3}; /* column-major */
  const double complex rb[4] = {0.1, 0.05 * I, -0.05 * I, 0.9};
  double complex m[16];
  for (int ia = 0; ia < 2; ia++)
    for (int ib = 0; ib < 2; ib++)
      for (int ja = 0; ja < 2; ja++)
        for (int jb = 0; jb < 2; jb++) m[4 * (2 * ja + jb) + 2 * ia + ib] = ra[2 * ja + ia] * rb[2 * jb + ib];
  create_full_dm(&rho);
  put(rho, m, 4);
  create_dm(&red, 2);
  partial_trace_over(rho, red, 1, b);
  show("over_b", red);
  destroy_dm(red);
  create_dm(&red, 2);
  partial_trace_keep(rho, red, 1, b);
  show("keep_b", red);
  destroy_dm(red);
  /* partial_trace_over_one: a is the middle factor of 1 x 2 x 2, b of 2 x 2 x 1 */
  create_dm(&red, 2);
  partial_trace_over_one(rho, red, 1, 2, 2, 4);
  show("one_a", red);
  destroy_dm(red);
  create_dm(&red, 2);
  partial_trace_over_one(rho, red, 2, 2, 1, 4);
  show("one_b", red);
  destroy_dm(red);
  create_full_dm(&bell);
  werner(bell, 1.0);
  for (int k = 0; k < 2; k++) {
    const double p = k ? 0.2 : 0.8;
    double conc = -9, fid = -9;
    create_full_dm(&w);
    werner(w, p);
    get_bipartite_concurrence(w, &conc);
    get_fidelity(w, bell, &fid);
    printf("werner %.2f %.15e %.15e\n", p, conc, fid);
    destroy_dm(w);
  }
  {
    /* cos t |01> + e^{i f} sin t |10>, t = 0.3, f = 0.7: concurrence sin 2t */
    double complex psi[4] = {0, cos(0.3), cexp(0.7 * I) * sin(0.3), 0}, pm[16];
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) pm[4 * c + r] = psi[r] * conj(psi[c]);
    double conc = -9;
    create_full_dm(&w);
    put(w, pm, 4);
    get_bipartite_concurrence(w, &conc);
    printf("pure %.15e\n", conc);
    destroy_dm(w);
  }
  {
    /* measure_dm with the projector n of a on a Werner state (p = 0.8) */
    create_full_dm(&w);
    werner(w, 0.8);
    measure_dm(w, a->n);
    printf("measure");
    for (int k = 0; k < 16; k++) {
      PetscScalar v;
      get_dm_element(w, k % 4, k / 4, &v);
      printf(" %.15e %.15e", creal(v), cimag(v));
    }
    printf("\n");
    destroy_dm(w);
  }
  /* sparse printers: the Bell state (4 entries of 1/2) and sigma_x of a as a
   * matrix (2 entries of 1 per row block: 4 entries) */
  printf("sparse_begin\n");
  print_dm_sparse(bell, 4);
  printf("sparse_end\n");
  print_dm_sparse_to_file(bell, 4, "dm_sparse.txt");
  {
    Mat sx = NULL;
    combine_ops_to_mat(&sx, total_levels, 2, a, a->dag); /* a a^dag = |0><0| (x) I */
    print_mat_sparse_to_file(sx, "mat_sparse.txt");
    qh_mat_destroy(sx);
  }
  destroy_dm(bell);
  destroy_dm(rho);
  QuaC_finalize();
  return 0;
}
