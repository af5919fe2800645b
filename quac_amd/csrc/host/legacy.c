/* legacy.c -- QuaC's original global-operator API (src/operators.c,
 * src/dm_utilities.c, src/solver.c) over one process-wide qsystem.
 *
 * The reference builds its global superoperator full_A (and the Schrodinger
 * ham_A) term by term with the same row rules as the qsystem path: add_to_ham
 * adds -i a (I (x) H) + i a (H^T (x) I) (operators.c:290-337), add_lin adds the
 * Lindblad dissipator of one operator (operators.c:679-719), the mult2/mult3
 * forms the term of the operator product, a VEC pair being |p1><p2|
 * (operators.c:405-488, 601-663, 743-821).  Here every such call records the
 * same term in the process-wide qsystem, and construct_matrix assembles it on
 * the device before the first state is made or evolved, and again when terms
 * were added since.  The legacy time_step runs TSRK3BS with PETSc's default
 * tolerances (solver.c:236-514 sets none), so unless the options say otherwise
 * the adaptive step uses atol = rtol = 1e-4 and no step-size bounds; the final
 * step is matched to time_max (the reference steps over it). */
#include <complex.h>
#include <math.h>
#include <string.h>

#include "quac_internal.h"
#include "quac_legacy.h"

int nid = 0;
int np = 1;
PetscInt total_levels = 1;
int num_subsystems = 0;
operator subsystem_list[MAX_SUB];

static qsystem g_sys = NULL;
static int g_dirty = 0;       /* terms added since the last assembly */
static int g_stiff = 0;       /* add_to_ham_stiff* used (refused by time_step) */
static PetscErrorCode (*g_mon)(TS, PetscInt, PetscReal, Vec, void *) = NULL;
static void *g_monctx = NULL;

/* states made by create_full_dm / create_dm: the Vec handle and its qvec */
#define LEGACY_MAXSTATE 256
static struct {
  Vec v;
  qvec q;
} g_state[LEGACY_MAXSTATE];

/* time-dependent coefficient functions of add_to_ham_time_dep (double f(double)) */
#define LEGACY_MAXTD 100
static struct {
  double (*f)(double);
} g_td[LEGACY_MAXTD];
static int g_ntd = 0;

qsystem quac_legacy_system(void) { return g_sys; }

/* operators.c:996-1046: the first operator creates the system */
static int ensure_system(void) {
  if (g_sys) return 0;
  initialize_system(&g_sys);
  if (!g_sys) {
    quac_fail(9, "cannot create the system");
    return -1;
  }
  total_levels = 1;
  num_subsystems = 0;
  return 0;
}

static int ensure_terms_allowed(void) {
  if (g_sys) return 0;
  quac_fail(0, "You need to create operators before you add terms!");
  return -1;
}

/* assemble (again) when terms were added since the last assembly */
static int ensure_matrix(void) {
  if (!g_sys) {
    quac_fail(0, "You need to create operators and add terms first!");
    return -1;
  }
  if (g_sys->mat_allocated == 1 && !g_dirty) return 0;
  construct_matrix(g_sys);
  if (g_sys->mat_allocated != 1) return -1; /* construct_matrix recorded the error */
  g_dirty = 0;
  return 0;
}

static int remember(Vec v, qvec q) {
  for (int i = 0; i < LEGACY_MAXSTATE; i++)
    if (!g_state[i].v) {
      g_state[i].v = v;
      g_state[i].q = q;
      return 0;
    }
  quac_fail(9, "more than %d legacy states alive", LEGACY_MAXSTATE);
  return -1;
}

static qvec lookup(Vec v) {
  for (int i = 0; i < LEGACY_MAXSTATE; i++)
    if (g_state[i].v == v && v) return g_state[i].q;
  return NULL;
}

/* ---------------------------------------------------------------- operators */
void create_op(int number_of_levels, operator *new_op) {
  if (ensure_system()) return;
  if (num_subsystems == MAX_SUB) QUAC_FAIL(0, "More than %d subsystems", MAX_SUB);
  create_op_sys(g_sys, number_of_levels, new_op);
  if (!*new_op) return;
  subsystem_list[num_subsystems++] = *new_op;
  total_levels = g_sys->total_levels;
}

void create_vec(int number_of_levels, vec_op *new_vec) {
  if (ensure_system()) return;
  if (num_subsystems == MAX_SUB) QUAC_FAIL(0, "More than %d subsystems", MAX_SUB);
  create_vec_op_sys(g_sys, number_of_levels, new_vec);
  if (!*new_vec) return;
  subsystem_list[num_subsystems++] = (*new_vec)[0];
  total_levels = g_sys->total_levels;
}

void destroy_op(operator *op) { destroy_op_sys(op); }
void destroy_vec(vec_op *op) { destroy_vec_op_sys(op); }

/* operators.c:923-994: the total_levels x total_levels matrix of the product
 * op1 op2 ... (subsystems not named act as the identity; no operators: the
 * identity), as a device CSR the caller destroys with qh_mat_destroy */
void combine_ops_to_mat(Mat *matrix_out, PetscInt tot_levels, int number_of_ops, ...) {
  *matrix_out = NULL;
  if (!g_sys || num_subsystems == 0) QUAC_FAIL(0, "You need to create operators before combine_ops_to_mat!");
  if (tot_levels != total_levels) QUAC_FAIL(9, "combine_ops_to_mat: tot_levels must be total_levels");
  if (number_of_ops < 0 || number_of_ops > 64) QUAC_FAIL(9, "combine_ops_to_mat: 0..64 operators");
  operator ops[64];
  va_list ap;
  va_start(ap, number_of_ops);
  for (int i = 0; i < number_of_ops; i++) ops[i] = va_arg(ap, operator);
  va_end(ap);
  if (number_of_ops == 0) {
    ops[0] = subsystem_list[0]->my_op_type == VEC ? subsystem_list[0]->vec_op_list[0]->eye : subsystem_list[0]->eye;
    number_of_ops = 1;
  }
  construct_op_matrix_wf_list(g_sys, 1.0, matrix_out, number_of_ops, ops);
}

void no_lindblad_terms(void) {} /* operators.c:278-280: a hint; the solver follows the terms */

void print_dense_ham(void) {} /* operators.c:39-48: the dense text dump is not produced */

void set_initial_pop(operator op, double initial_pop) { op->initial_pop = initial_pop; }

/* operators.c:1048-1096: VEC operators multiply only with VEC operators of the
 * same subsystem */
static int check_pair(operator op1, operator op2) {
  const char *why = NULL;
  if ((op1->my_op_type == VEC) != (op2->my_op_type == VEC))
    why = "Multiplying a VEC_OP and a regular OP does not make sense!";
  else if (op1->my_op_type == VEC && op1->n_before != op2->n_before)
    why = "Multiplying two VEC_OPs from different subspaces does not make sense!";
  if (!why) return 0;
  quac_fail(0, "%s", why);
  return -1;
}

void add_to_ham(PetscScalar a, operator op) {
  if (ensure_terms_allowed()) return;
  if (cabs(a) == 0) return; /* operators.c:297: zero terms are skipped */
  add_ham_term(g_sys, a, 1, op);
  g_dirty = 1;
}

void add_to_ham_stiff(PetscScalar a, operator op) {
  g_stiff = 1;
  add_to_ham(a, op);
}

void add_to_ham_mult2(PetscScalar a, operator op1, operator op2) {
  if (ensure_terms_allowed() || check_pair(op1, op2)) return;
  add_ham_term(g_sys, a, 2, op1, op2);
  g_dirty = 1;
}

void add_to_ham_stiff_mult2(PetscScalar a, operator op1, operator op2) {
  g_stiff = 1;
  add_to_ham_mult2(a, op1, op2);
}

void add_to_ham_mult3(PetscScalar a, operator op1, operator op2, operator op3) {
  if (ensure_terms_allowed()) return;
  add_ham_term(g_sys, a, 3, op1, op2, op3);
  g_dirty = 1;
}

void add_lin(PetscScalar a, operator op) {
  if (ensure_terms_allowed()) return;
  add_lin_term(g_sys, a, 1, op);
  g_dirty = 1;
}

void add_lin_mult2(PetscScalar a, operator op1, operator op2) {
  if (ensure_terms_allowed() || check_pair(op1, op2)) return;
  add_lin_term(g_sys, a, 2, op1, op2);
  g_dirty = 1;
}

void add_lin_mat(PetscScalar a, Mat m) {
  (void)a;
  (void)m;
  QUAC_FAIL(9, "add_lin_mat: Lindblad terms from an arbitrary matrix are not supported (use operator terms)");
}

static void td_value(PetscReal t, void *ctx, PetscReal *re, PetscReal *im) {
  *re = g_td[(intptr_t)ctx].f(t);
  *im = 0.0;
}

/* operators.c:253-276: H(t) += f(t) op1 op2 ... */
void add_to_ham_time_dep(double (*time_dep_func)(double), int num_ops, ...) {
  if (ensure_terms_allowed()) return;
  if (g_ntd == LEGACY_MAXTD) QUAC_FAIL(9, "more than %d time-dependent terms", LEGACY_MAXTD);
  if (num_ops < 1 || num_ops > 64) QUAC_FAIL(9, "add_to_ham_time_dep: 1..64 operators");
  operator ops[64];
  va_list ap;
  va_start(ap, num_ops);
  for (int i = 0; i < num_ops; i++) ops[i] = va_arg(ap, operator);
  va_end(ap);
  g_td[g_ntd].f = time_dep_func;
  add_ham_term_time_dep_list(g_sys, 1.0, (void *)(intptr_t)g_ntd, td_value, num_ops, ops);
  g_ntd++;
  g_dirty = 1;
}

/* ------------------------------------------------------------ density matrices */
/* dm_utilities.c:635-655: a zero state of the system's dimension (a density
 * matrix when there are Lindblad terms, else a wavefunction) */
void create_full_dm(Vec *new_dm) {
  *new_dm = NULL;
  if (!g_sys) QUAC_FAIL(0, "You need to create operators before you create a density matrix!");
  qvec q = NULL;
  create_qvec_sys(g_sys, &q);
  if (!q) return;
  if (remember(q->data, q)) {
    destroy_qvec(&q);
    return;
  }
  *new_dm = q->data;
}

/* dm_utilities.c:616-623: a zero density matrix of size x size */
void create_dm(Vec *new_dm, PetscInt size) {
  *new_dm = NULL;
  qvec q = NULL;
  create_arb_qvec(&q, size * size, DENSITY_MATRIX);
  if (!q) return;
  if (remember(q->data, q)) {
    destroy_qvec(&q);
    return;
  }
  *new_dm = q->data;
}

void destroy_dm(Vec dm) {
  for (int i = 0; i < LEGACY_MAXSTATE; i++)
    if (g_state[i].v == dm && dm) {
      destroy_qvec(&g_state[i].q);
      g_state[i].v = NULL;
      g_state[i].q = NULL;
      return;
    }
}

#define LOOKUP_OR_FAIL(q, v)                                                       \
  qvec q = lookup(v);                                                              \
  if (!q) QUAC_FAIL(9, "not a state made by create_full_dm / create_dm")

/* population of level l of subsystem s in the initial product state */
static double sub_pop(operator s, PetscInt l) {
  if (s->my_op_type != VEC) return l == (PetscInt)s->initial_pop ? 1.0 : 0.0;
  double tot = 0.0;
  for (PetscInt j = 0; j < s->my_levels; j++) tot += s->vec_op_list[j]->initial_pop;
  if (tot == 0.0) return l == 0 ? 1.0 : 0.0; /* dm_utilities.c:767-773 */
  double p = 0.0;
  for (PetscInt j = 0; j < s->my_levels; j++)
    if (s->vec_op_list[j]->position == l) p += s->vec_op_list[j]->initial_pop;
  return p / tot;
}

/* dm_utilities.c:665-804: the product state of the subsystems' initial
 * populations (ladder: the level initial_pop; VEC: the populations of its
 * levels, normalised).  Without VEC subsystems one entry is set; otherwise
 * the whole density matrix is written (diagonal, off-diagonals zero). */
void set_dm_from_initial_pop(Vec x) {
  LOOKUP_OR_FAIL(q, x);
  const int dm = q->my_type == DENSITY_MATRIX;
  int simple = 1;
  for (int i = 0; i < num_subsystems; i++)
    if (subsystem_list[i]->my_op_type == VEC) simple = 0;
  if (simple) {
    PetscInt row = 0;
    for (int i = 0; i < num_subsystems; i++) {
      const PetscInt n_after = total_levels / (subsystem_list[i]->my_levels * subsystem_list[i]->n_before);
      row += (PetscInt)subsystem_list[i]->initial_pop * n_after;
    }
    const PetscInt loc = dm ? total_levels * row + row : row;
    double re = 0.0, im = 0.0;
    QUAC_HIPCHK(qh_vec_get_entry(x, loc, &re, &im));
    QUAC_HIPCHK(qh_vec_add_entry(x, loc, 1.0 - re, -im)); /* INSERT_VALUES */
    return;
  }
  if (!dm) QUAC_FAIL(9, "set_dm_from_initial_pop: VEC populations need a density matrix (Lindblad terms)");
  const PetscInt N = total_levels;
  double *h = (double *)calloc((size_t)(2 * N * N), sizeof(double));
  if (!h) QUAC_FAIL(9, "out of host memory");
  for (PetscInt r = 0; r < N; r++) {
    double p = 1.0;
    for (int i = 0; i < num_subsystems && p != 0.0; i++) {
      operator s = subsystem_list[i];
      const PetscInt n_after = N / (s->my_levels * s->n_before);
      p *= sub_pop(s, (r / n_after) % s->my_levels);
    }
    h[2 * (N * r + r)] = p;
  }
  const int rc = qh_vec_set(x, h);
  free(h);
  if (rc) QUAC_FAIL(9, "set_dm_from_initial_pop: %s", qh_last_error());
}

/* dm_utilities.c:965-981 (location sqrt(size) col + row); *val is written */
void get_dm_element(Vec dm, PetscInt row, PetscInt col, PetscScalar *val) {
  LOOKUP_OR_FAIL(q, dm);
  get_dm_element_qvec(q, row, col, val);
}

/* dm_utilities.c:999-1016: the same element (one process holds every entry) */
void get_dm_element_local(Vec dm, PetscInt row, PetscInt col, PetscScalar *val) { get_dm_element(dm, row, col, val); }

/* dm_utilities.c:1049-1061 */
void add_value_to_dm(Vec dm, PetscInt row, PetscInt col, PetscScalar val) {
  LOOKUP_OR_FAIL(q, dm);
  const PetscInt n = q->total_levels;
  QUAC_HIPCHK(qh_vec_add_entry(dm, n * col + row, creal(val), cimag(val)));
}

void assemble_dm(Vec dm) { (void)dm; } /* values are in device memory as soon as they are added */

/* dm_utilities.c:1127-1143: one population per ladder subsystem, one per level
 * of a VEC subsystem */
int get_num_populations(void) {
  int n = 0;
  for (int i = 0; i < num_subsystems; i++) n += subsystem_list[i]->my_op_type == VEC ? (int)subsystem_list[i]->my_levels : 1;
  return n;
}

/* dm_utilities.c:1157-1325: <n> of each ladder subsystem and the population of
 * each level of each VEC subsystem, from the diagonal (|psi_i|^2 for a
 * wavefunction); (*populations) is the caller's array of get_num_populations() */
void get_populations(Vec x, double **populations) {
  LOOKUP_OR_FAIL(q, x);
  PetscInt nl = 0;
  PetscReal *probs = NULL, *vars = NULL;
  get_bitstring_probs(q, &nl, &probs, &vars);
  if (!probs) return;
  int *first = (int *)malloc(sizeof(int) * (size_t)(num_subsystems + 1));
  int num_pop = 0;
  for (int i = 0; i < num_subsystems; i++) {
    first[i] = num_pop;
    num_pop += subsystem_list[i]->my_op_type == VEC ? (int)subsystem_list[i]->my_levels : 1;
  }
  for (int i = 0; i < num_pop; i++) (*populations)[i] = 0.0;
  for (PetscInt r = 0; r < nl; r++)
    for (int j = 0; j < num_subsystems; j++) {
      operator s = subsystem_list[j];
      const PetscInt n_after = total_levels / (s->my_levels * s->n_before);
      const PetscInt cur = (r / n_after) % s->my_levels;
      if (s->my_op_type == VEC) (*populations)[first[j] + cur] += probs[r];
      else (*populations)[first[j]] += (double)cur * probs[r];
    }
  free(first);
  free(probs);
  free(vars);
}

/* dm_utilities.c:1352-1469: Tr(O rho) of the operator product (<psi|O|psi> for
 * a wavefunction) */
void get_expectation_value(Vec rho, PetscScalar *trace_val, int number_of_ops, ...) {
  LOOKUP_OR_FAIL(q, rho);
  if (number_of_ops < 1 || number_of_ops > 64) QUAC_FAIL(9, "get_expectation_value: 1..64 operators");
  operator ops[64];
  va_list ap;
  va_start(ap, number_of_ops);
  for (int i = 0; i < number_of_ops; i++) ops[i] = va_arg(ap, operator);
  va_end(ap);
  get_expectation_value_qvec_list(q, trace_val, number_of_ops, ops);
}

/* dm_utilities.c:1766-1786 */
void trace_dm(PetscScalar *trace_val, Vec dm) {
  LOOKUP_OR_FAIL(q, dm);
  get_trace_qvec(q, trace_val);
}

/* quantum_gates.c:755-760: run the circuit during the next time_step, its gate
 * times counted from `time` */
void start_circuit_at_time(circuit *circ, PetscReal time) {
  if (!g_sys) QUAC_FAIL(0, "You need to create operators before you start a circuit!");
  apply_circuit_to_sys(g_sys, circ, time);
}

/* index of the subsystem an operator (or any level of a VEC subsystem) acts on */
static int subsystem_of(operator op) {
  for (int i = 0; i < num_subsystems; i++)
    if (subsystem_list[i]->n_before == op->n_before && subsystem_list[i]->my_levels == op->my_levels) return i;
  return -1;
}

/* the reduced density matrix over `traced` (flags per subsystem) into pt */
static void ptrace_into(Vec full_dm, Vec ptraced_dm, const int *traced) {
  LOOKUP_OR_FAIL(q, full_dm);
  LOOKUP_OR_FAIL(p, ptraced_dm);
  PetscInt list[MAX_SUB], n = 0;
  for (int i = 0; i < num_subsystems; i++)
    if (traced[i]) list[n++] = i;
  qvec red = NULL;
  ptrace_over_list_qvec(q, n, list, &red);
  if (!red) return;
  if (red->n != p->n) {
    destroy_qvec(&red);
    QUAC_FAIL(0, "the partial trace has %lld entries but the target density matrix %lld", (long long)red->n,
              (long long)p->n);
  }
  const int rc = qh_vec_copy(red->data, ptraced_dm);
  destroy_qvec(&red);
  if (rc) QUAC_FAIL(9, "partial trace: %s", qh_last_error());
}

/* dm_utilities.c:150-260: trace over the listed subsystems */
void partial_trace_over(Vec full_dm, Vec ptraced_dm, int number_of_ops, ...) {
  int traced[MAX_SUB] = {0};
  va_list ap;
  va_start(ap, number_of_ops);
  for (int i = 0; i < number_of_ops; i++) {
    const int k = subsystem_of(va_arg(ap, operator));
    if (k >= 0) traced[k] = 1;
  }
  va_end(ap);
  ptrace_into(full_dm, ptraced_dm, traced);
}

/* dm_utilities.c:468-603: keep the listed subsystems, trace over the others */
void partial_trace_keep(Vec full_dm, Vec ptraced_dm, int number_of_ops, ...) {
  int traced[MAX_SUB];
  for (int i = 0; i < MAX_SUB; i++) traced[i] = 1;
  va_list ap;
  va_start(ap, number_of_ops);
  for (int i = 0; i < number_of_ops; i++) {
    const int k = subsystem_of(va_arg(ap, operator));
    if (k >= 0) traced[k] = 0;
  }
  va_end(ap);
  ptrace_into(full_dm, ptraced_dm, traced);
}

/* dm_utilities.c:1078-1117: trace over the middle factor of a space
 * nbef x nop x naf (cur_levels = nbef nop naf), ADDING into ptraced_dm
 * (nbef naf)^2.  Entry (r, c) of either matrix sits at dim r + c here, as in
 * the reference; the partial trace commutes with the transpose, so this is the
 * partial trace in either reading. */
void partial_trace_over_one(Vec full_dm, Vec ptraced_dm, PetscInt nbef, PetscInt nop, PetscInt naf,
                            PetscInt cur_levels) {
  LOOKUP_OR_FAIL(q, full_dm);
  LOOKUP_OR_FAIL(p, ptraced_dm);
  const PetscInt ns = nbef * naf;
  if (nbef < 1 || nop < 1 || naf < 1 || cur_levels != nbef * nop * naf || q->n != cur_levels * cur_levels ||
      p->n != ns * ns)
    QUAC_FAIL(0, "partial_trace_over_one: the sizes do not match the density matrices");
  double complex *f = (double complex *)malloc(sizeof(double complex) * (size_t)q->n);
  double complex *r = (double complex *)malloc(sizeof(double complex) * (size_t)p->n);
  int rc = !f || !r;
  if (!rc) rc = qh_vec_get(full_dm, (double *)f) || qh_vec_get(ptraced_dm, (double *)r);
  if (!rc) {
    for (PetscInt ib = 0; ib < nbef; ib++)
      for (PetscInt jb = 0; jb < nbef; jb++)
        for (PetscInt ia = 0; ia < naf; ia++)
          for (PetscInt ja = 0; ja < naf; ja++)
            for (PetscInt k = 0; k < nop; k++)
              r[ns * (naf * ib + ia) + naf * jb + ja] +=
                  f[cur_levels * (naf * nop * ib + naf * k + ia) + naf * nop * jb + naf * k + ja];
    rc = qh_vec_set(ptraced_dm, (const double *)r);
  }
  free(f);
  free(r);
  if (rc) QUAC_FAIL(9, "partial_trace_over_one: %s", f && r ? qh_last_error() : "out of host memory");
}

/* dm_utilities.c:1618-1722: Tr sqrt(sqrt(rho) sigma sqrt(rho)), positive
 * eigenvalues only -- not squared, unlike get_fidelity_qvec, whose value is the
 * square of this one */
void get_fidelity(Vec dm, Vec dm_r, double *fidelity) {
  LOOKUP_OR_FAIL(a, dm);
  LOOKUP_OR_FAIL(b, dm_r);
  if (a->n != b->n) QUAC_FAIL(0, "get_fidelity: the density matrices differ in size");
  PetscReal f2 = -1.0, var = 0.0;
  get_fidelity_qvec(a, b, &f2, &var);
  if (f2 < 0.0) return;
  *fidelity = sqrt(f2);
}

/* dm_utilities.c:1487-1602: concurrence of a two-qubit density matrix */
void get_bipartite_concurrence(Vec dm, double *concurrence) {
  LOOKUP_OR_FAIL(q, dm);
  if (q->n != 16) QUAC_FAIL(0, "The input density matrix is not 4x4! Concurrence cannot be calculated.");
  double complex rho[16];
  QUAC_HIPCHK(qh_vec_get(dm, (double *)rho));
  quac_concurrence_2q(rho, concurrence);
}

/* dm_utilities.c:270-308: rho <- A rho A^dag / Tr(A rho A^dag) for the operator A
 * (its full total_levels-dimensional matrix); rho is left unchanged when that
 * trace is 0.  Host arithmetic over A's CSR: O(N^2 nnz/row). */
void measure_dm(Vec dm, operator op) {
  LOOKUP_OR_FAIL(q, dm);
  if (q->my_type != DENSITY_MATRIX || q->total_levels != total_levels)
    QUAC_FAIL(0, "measure_dm: a full density matrix of the system is required");
  const PetscInt N = total_levels;
  Mat m = NULL;
  combine_ops_to_mat(&m, total_levels, 1, op);
  if (!m) return;
  int64_t rows = 0, nnz = 0;
  qh_mat_info(m, &rows, &nnz);
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(rows + 1));
  int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
  double complex *val = (double complex *)malloc(sizeof(double complex) * (size_t)(nnz > 0 ? nnz : 1));
  double complex *rho = (double complex *)malloc(sizeof(double complex) * (size_t)(N * N));
  double complex *T = (double complex *)calloc((size_t)(N * N), sizeof(double complex));
  double complex *R = (double complex *)calloc((size_t)(N * N), sizeof(double complex));
  int rc = !rp || !col || !val || !rho || !T || !R || rows != N;
  if (!rc) rc = qh_mat_export(m, rp, col, (double *)val) || qh_vec_get(dm, (double *)rho);
  qh_mat_destroy(m);
  if (!rc) {
    /* T = A rho, then R = T A^dag: R(i, j) = sum_k T(i, k) conj(A(j, k)); column-major */
    for (PetscInt i = 0; i < N; i++)
      for (int64_t e = rp[i]; e < rp[i + 1]; e++)
        for (PetscInt c = 0; c < N; c++) T[N * c + i] += val[e] * rho[N * c + col[e]];
    for (PetscInt j = 0; j < N; j++)
      for (int64_t e = rp[j]; e < rp[j + 1]; e++)
        for (PetscInt i = 0; i < N; i++) R[N * j + i] += T[N * col[e] + i] * conj(val[e]);
    double complex tr = 0.0;
    for (PetscInt i = 0; i < N; i++) tr += R[N * i + i];
    if (tr != 0.0) {
      for (PetscInt k = 0; k < N * N; k++) R[k] /= tr;
      rc = qh_vec_set(dm, (const double *)R);
    }
  }
  free(rp);
  free(col);
  free(val);
  free(rho);
  free(T);
  free(R);
  if (rc) QUAC_FAIL(9, "measure_dm: %s", rows != N ? "operator size mismatch" : qh_last_error());
}

/* the density matrix as host complex128 after checking h_dim (caller frees) */
static double complex *dm_block(Vec rho, int h_dim, const char *fn) {
  qvec q = lookup(rho);
  if (!q) return quac_fail(9, "%s: not a state made by create_full_dm / create_dm", fn), NULL;
  if (q->my_type != DENSITY_MATRIX || h_dim < 0 || h_dim > q->total_levels)
    return quac_fail(0, "%s: h_dim %d does not fit the density matrix", fn, h_dim), NULL;
  double complex *h = (double complex *)malloc(sizeof(double complex) * (size_t)q->n);
  if (!h) return quac_fail(9, "out of host memory"), NULL;
  if (qh_vec_get(rho, (double *)h)) {
    free(h);
    return quac_fail(9, "%s: %s", fn, qh_last_error()), NULL;
  }
  return h;
}

/* element (i, j) of an N x N density matrix stored column-major */
#define DM_AT(h, N, i, j) ((h)[(size_t)(N) * (size_t)(j) + (size_t)(i)])

/* dm_utilities.c:13-25: the leading h_dim x h_dim block of the density matrix */
void print_dm(Vec rho, int h_dim) {
  double complex *h = dm_block(rho, h_dim, "print_dm");
  if (!h) return;
  const PetscInt N = lookup(rho)->total_levels;
  for (int i = 0; i < h_dim; i++) {
    for (int j = 0; j < h_dim; j++) printf("%4.3e + %4.3ei ", creal(DM_AT(h, N, i, j)), cimag(DM_AT(h, N, i, j)));
    printf("\n");
  }
  printf("\n");
  free(h);
}

/* dm_utilities.c:32-43 / 51-66: "i j re im" for the entries above 1e-10 in modulus */
static void dm_sparse(FILE *fp, Vec rho, int h_dim, const char *fn) {
  double complex *h = dm_block(rho, h_dim, fn);
  if (!h) return;
  const PetscInt N = lookup(rho)->total_levels;
  for (int i = 0; i < h_dim; i++)
    for (int j = 0; j < h_dim; j++)
      if (cabs(DM_AT(h, N, i, j)) > 1e-10)
        fprintf(fp, "%d %d %e %e\n", i, j, creal(DM_AT(h, N, i, j)), cimag(DM_AT(h, N, i, j)));
  free(h);
}

void print_dm_sparse(Vec rho, int h_dim) { dm_sparse(stdout, rho, h_dim, "print_dm_sparse"); }

void print_dm_sparse_to_file(Vec rho, int h_dim, char filename[]) {
  FILE *fp = fopen(filename, "w");
  if (!fp) QUAC_FAIL(1, "print_dm_sparse_to_file: cannot open %s", filename);
  dm_sparse(fp, rho, h_dim, "print_dm_sparse_to_file");
  fclose(fp);
}

/* dm_utilities.c:71-92 / 98-113: the stored entries of a matrix, row by row;
 * the file form drops entries of modulus <= 1e-10 and prints 24 digits */
static void mat_sparse(FILE *fp, Mat A, int to_file) {
  int64_t rows = 0, nnz = 0;
  if (qh_mat_info(A, &rows, &nnz)) QUAC_FAIL(9, "print_mat_sparse: %s", qh_last_error());
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(rows + 1));
  int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
  double *val = (double *)malloc(sizeof(double) * 2 * (size_t)(nnz > 0 ? nnz : 1));
  if (!rp || !col || !val || qh_mat_export(A, rp, col, val)) {
    free(rp);
    free(col);
    free(val);
    QUAC_FAIL(9, "print_mat_sparse: %s", rp && col && val ? qh_last_error() : "out of host memory");
  }
  for (int64_t r = 0; r < rows; r++)
    for (int64_t e = rp[r]; e < rp[r + 1]; e++) {
      const double re = val[2 * e], im = val[2 * e + 1];
      if (!to_file) fprintf(fp, "%d %d %e %e\n", (int)r, (int)col[e], re, im);
      else if (hypot(re, im) > 1e-10) fprintf(fp, "%d %d %12.24e %12.24e\n", (int)r, (int)col[e], re, im);
    }
  free(rp);
  free(col);
  free(val);
}

void print_mat_sparse(Mat A) { mat_sparse(stdout, A, 0); }

void print_mat_sparse_to_file(Mat A, char filename[]) {
  FILE *fp = fopen(filename, "w");
  if (!fp) QUAC_FAIL(1, "print_mat_sparse_to_file: cannot open %s", filename);
  mat_sparse(fp, A, 1);
  fclose(fp);
}

/* dm_utilities.c:120-132: the first h_dim entries of the vector */
void print_psi(Vec rho, int h_dim) {
  LOOKUP_OR_FAIL(q, rho);
  if (h_dim < 0 || h_dim > q->n) QUAC_FAIL(0, "print_psi: h_dim %d exceeds the vector", h_dim);
  double complex *h = (double complex *)malloc(sizeof(double complex) * (size_t)q->n);
  if (!h) QUAC_FAIL(9, "out of host memory");
  if (qh_vec_get(rho, (double *)h)) {
    free(h);
    QUAC_FAIL(9, "print_psi: %s", qh_last_error());
  }
  for (int i = 0; i < h_dim; i++) printf("%f + %f i\n", creal(h[i]), cimag(h[i]));
  printf("\n");
  free(h);
}

/* ------------------------------------------------------------------- solvers */
void set_ts_monitor(PetscErrorCode (*monitor)(TS, PetscInt, PetscReal, Vec, void *)) {
  g_mon = monitor;
  g_monctx = NULL;
}

void set_ts_monitor_ctx(PetscErrorCode (*monitor)(TS, PetscInt, PetscReal, Vec, void *), void *ctx) {
  g_mon = monitor;
  g_monctx = ctx;
}

/* set name to value for the duration of a call unless the caller set it */
typedef struct {
  const char *name;
  int was_set;
} opt_default;

static void push_default(opt_default *d, const char *name, const char *value) {
  d->name = name;
  d->was_set = quac_get_option(name) != NULL;
  if (!d->was_set) quac_set_option(name, value);
}

static void pop_default(const opt_default *d) {
  if (!d->was_set) quac_unset_option(d->name);
}

/* solver.c:236-514 */
void time_step(Vec x, PetscReal init_time, PetscReal time_max, PetscReal dt, PetscInt steps_max) {
  if (g_stiff)
    QUAC_FAIL(0, "%s", g_sys && g_sys->dm_equations ? "Lindblad-stiff solver untested." : "Stiff solver not implemented!");
  if (ensure_matrix()) return;
  LOOKUP_OR_FAIL(q, x);
  set_ts_monitor_sys(g_sys, g_mon, g_monctx);
  opt_default d[4];
  push_default(&d[0], "ts_atol", "1e-4");
  push_default(&d[1], "ts_rtol", "1e-4");
  push_default(&d[2], "ts_adapt_dt_min", "1e-20");
  push_default(&d[3], "ts_adapt_dt_max", "1e20");
  time_step_sys(g_sys, q, init_time, time_max, dt, steps_max);
  for (int i = 3; i >= 0; i--) pop_default(&d[i]);
}

/* solver.c:33-221: the stabilised Liouvillian solve of the qsystem path */
void steady_state(Vec x) {
  if (!g_sys || !g_sys->dm_equations)
    QUAC_FAIL(0, "Warning! Steady state not supported for Schrodinger.");
  if (ensure_matrix()) return;
  LOOKUP_OR_FAIL(q, x);
  steady_state_sys(g_sys, q);
}
