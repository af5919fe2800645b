/* qsystem.c -- qsystem object, operator factory, term registry, construct_matrix
 * and time_step_sys of the QuaC-compatible host layer.
 *
 * Follows src/qsystem.c (line numbers cited per function).  The matrix is built
 * on the device (qh_assemble) and every step runs in fused gfx950 kernels.
 */
#include <math.h>
#include <string.h>

#include "quac_internal.h"

#define NUM_INIT_ALLOC 150 /* qsystem.c:18 */

/* ----------------------------------------------------------- qsystem.c:16 */
void initialize_system(qsystem *qsys) {
  if (quac_ensure_device()) return;
  qsystem t = (qsystem)calloc(1, sizeof(struct qsystem));
  t->hspace_frozen = 0;
  t->dm_equations = PETSC_FALSE;
  t->mcwf_solver = PETSC_FALSE;
  t->time_step_called = PETSC_FALSE;
  t->total_levels = 1;
  t->dim = 1;
  t->min_time_dep = 1e-10;
  t->alloc_subsystems = NUM_INIT_ALLOC;
  t->subsystem_list = (operator *)malloc(NUM_INIT_ALLOC * sizeof(operator));
  t->alloc_time_dep = NUM_INIT_ALLOC;
  t->alloc_time_indep = NUM_INIT_ALLOC;
  t->time_indep = (mat_term *)calloc(NUM_INIT_ALLOC, sizeof(mat_term));
  t->time_dep = (mat_term *)calloc(NUM_INIT_ALLOC, sizeof(mat_term));
  t->np = 1;
  t->nid = 0;
  t->my_num = -1;
  t->Istart = -1;
  t->Iend = -1;
  t->num_circuits = -1;
  t->current_circuit = 0;
  *qsys = t;
}

/* ----------------------------------------------------------- qsystem.c:67 */
void destroy_system(qsystem *qsys) {
  if (!qsys || !*qsys) return;
  quac_mcwf_free(*qsys);
  clear_mat_terms_sys(*qsys);
  free((*qsys)->subsystem_list);
  free((*qsys)->time_dep);
  free((*qsys)->time_indep);
  if ((*qsys)->num_circuits > 0) free((*qsys)->circuit_list);
  free(*qsys);
  *qsys = NULL;
}

/* ----------------------------------------------------------- qsystem.c:93 */
static operator single_op(PetscInt total_levels, PetscInt levels, op_type t) {
  operator o = (operator)calloc(1, sizeof(struct operator));
  o->n_before = total_levels;
  o->my_levels = levels;
  o->my_op_type = t;
  o->position = -1;
  const double s = 1.0 / sqrt(2.0);
  if (t == SIGMA_Z) {
    o->evals[0] = 1;
    o->evals[1] = -1;
    o->evecs[0][0] = -1;
    o->evecs[1][1] = -1;
  } else if (t == SIGMA_Y) {
    o->evals[0] = -1;
    o->evals[1] = 1;
    o->evecs[0][0] = -s;
    o->evecs[0][1] = s * PETSC_i;
    o->evecs[1][0] = -s;
    o->evecs[1][1] = -s * PETSC_i;
  } else if (t == SIGMA_X) {
    o->evals[0] = -1;
    o->evals[1] = 1;
    o->evecs[0][0] = -s;
    o->evecs[0][1] = s;
    o->evecs[1][0] = s;
    o->evecs[1][1] = s;
  }
  return o;
}

static void push_subsystem(qsystem sys, operator op) {
  if (sys->num_subsystems == sys->alloc_subsystems) {
    sys->alloc_subsystems *= 2;
    sys->subsystem_list = (operator *)realloc(sys->subsystem_list, sys->alloc_subsystems * sizeof(operator));
  }
  op->pos_in_sys_hspace = sys->num_subsystems;
  sys->subsystem_list[sys->num_subsystems++] = op;
}

/* ---------------------------------------------------------- qsystem.c:153 */
void create_op_sys(qsystem sys, PetscInt levels, operator *new_op) {
  if (sys->hspace_frozen)
    QUAC_FAIL(0, "You cannot add more operators after the creating a dm or constructing the matrix!");
  if (levels < 1) QUAC_FAIL(0, "number of levels must be positive");
  operator lo = single_op(sys->total_levels, levels, LOWER);
  operator ra = single_op(sys->total_levels, levels, RAISE);
  ra->dag = lo;
  lo->dag = ra;
  lo->n = single_op(sys->total_levels, levels, NUMBER);
  lo->n->dag = lo->n;
  lo->eye = single_op(sys->total_levels, levels, IDENTITY);
  lo->eye->dag = lo->eye;
  lo->sig_x = single_op(sys->total_levels, levels, SIGMA_X);
  lo->sig_x->dag = lo->sig_x;
  lo->sig_z = single_op(sys->total_levels, levels, SIGMA_Z);
  lo->sig_z->dag = lo->sig_z;
  lo->sig_y = single_op(sys->total_levels, levels, SIGMA_Y);
  lo->sig_y->dag = lo->sig_y;
  lo->other = single_op(sys->total_levels, levels, OTHER);
  lo->other->dag = lo->other;
  sys->total_levels *= levels;
  push_subsystem(sys, lo);
  *new_op = lo;
}

/* ---------------------------------------------------------- qsystem.c:338 */
void destroy_op_sys(operator *op) {
  if (!op || !*op) return;
  free((*op)->dag);
  free((*op)->n);
  free((*op)->eye);
  free((*op)->sig_x);
  free((*op)->sig_z);
  free((*op)->sig_y);
  free((*op)->other);
  free(*op);
  *op = NULL;
}

/* ---------------------------------------------------------- qsystem.c:265 */
void create_vec_op_sys(qsystem sys, PetscInt levels, vec_op *new_vec) {
  if (sys->hspace_frozen)
    QUAC_FAIL(0, "You cannot add more vec_ops after the creating a dm or constructing the matrix!");
  *new_vec = (vec_op)malloc(levels * sizeof(operator));
  for (PetscInt i = 0; i < levels; i++) {
    operator v = single_op(sys->total_levels, levels, VEC);
    v->position = i;
    v->eye = single_op(sys->total_levels, levels, IDENTITY);
    v->eye->dag = v->eye;
    v->sig_x = single_op(sys->total_levels, levels, SIGMA_X);
    v->sig_x->dag = v->sig_x;
    v->sig_z = single_op(sys->total_levels, levels, SIGMA_Z);
    v->sig_z->dag = v->sig_z;
    v->sig_y = single_op(sys->total_levels, levels, SIGMA_Y);
    v->sig_y->dag = v->sig_y;
    v->other = single_op(sys->total_levels, levels, OTHER);
    v->other->dag = v->other;
    (*new_vec)[i] = v;
  }
  (*new_vec)[0]->vec_op_list = *new_vec;
  sys->total_levels *= levels;
  push_subsystem(sys, (*new_vec)[0]);
}

void destroy_vec_op_sys(vec_op *vop) {
  if (!vop || !*vop) return;
  PetscInt L = (*vop)[0]->my_levels;
  for (PetscInt i = 0; i < L; i++) {
    operator v = (*vop)[i];
    free(v->eye);
    free(v->sig_x);
    free(v->sig_z);
    free(v->sig_y);
    free(v->other);
    free(v);
  }
  free(*vop);
  *vop = NULL;
}

/* ---------------------------------------------------------- qsystem.c:358 */
void use_mcwf_solver(qsystem qsys, PetscInt num_tot_trajs, PetscInt seed) {
  if (qsys->hspace_frozen == PETSC_TRUE)
    QUAC_FAIL(9, "use_mcwf_solver must be called before construct_matrix or create_qvec_sys!");
  if (qsys->dm_equations == PETSC_FALSE) QUAC_FAIL(9, "use_mcwf_solver must be used with a Lindblad term!");
  if (num_tot_trajs < 2) QUAC_FAIL(9, "use_mcwf_solver must be called with more than 2 num_tot_trajs!");
  qsys->hspace_frozen = PETSC_TRUE;
  qsys->mcwf_solver = PETSC_TRUE;
  qsys->seed = seed;
  qsys->num_tot_trajs = num_tot_trajs;
  qsys->num_local_trajs = num_tot_trajs;
  quac_rng_seed(qsys, (uint64_t)seed);
}

/* --------------------------------------------------------- term registry */
static void add_term(mat_term *list, PetscInt *count, PetscInt alloc, mat_term_type type, PetscScalar a,
                     PetscInt num_ops, operator *ops, void *ctx, PetscScalar (*f)(PetscReal, void *)) {
  if (*count >= alloc) QUAC_FAIL(9, "Asking for more terms than were allocated!");
  if (num_ops < 0) QUAC_FAIL(9, "negative operator count");
  mat_term *t = &list[*count];
  t->my_term_type = type;
  t->a = a;
  t->num_ops = num_ops;
  t->ops = (operator *)malloc((num_ops > 0 ? num_ops : 1) * sizeof(operator));
  for (PetscInt i = 0; i < num_ops; i++) t->ops[i] = ops[i];
  t->ctx = ctx;
  t->time_dep_func = f;
  t->time_dep_func_ri = NULL;
  t->mat_A = NULL;
  (*count)++;
}

static operator *collect(PetscInt n, va_list ap) {
  operator *ops = (operator *)malloc((n > 0 ? n : 1) * sizeof(operator));
  for (PetscInt i = 0; i < n; i++) ops[i] = va_arg(ap, operator);
  return ops;
}

/* qsystem.c:407-434 */
void add_ham_term(qsystem sys, PetscScalar a, PetscInt num_ops, ...) {
  va_list ap;
  va_start(ap, num_ops);
  operator *ops = collect(num_ops, ap);
  va_end(ap);
  add_term(sys->time_indep, &sys->num_time_indep, sys->alloc_time_indep, HAM, a, num_ops, ops, NULL, NULL);
  free(ops);
}

void add_ham_term_list(qsystem sys, PetscScalar a, PetscInt num_ops, operator *ops) {
  add_term(sys->time_indep, &sys->num_time_indep, sys->alloc_time_indep, HAM, a, num_ops, ops, NULL, NULL);
}

/* qsystem.c:436-462 */
void add_lin_term(qsystem sys, PetscScalar a, PetscInt num_ops, ...) {
  va_list ap;
  va_start(ap, num_ops);
  operator *ops = collect(num_ops, ap);
  va_end(ap);
  sys->dm_equations = PETSC_TRUE;
  add_term(sys->time_indep, &sys->num_time_indep, sys->alloc_time_indep, LINDBLAD, a, num_ops, ops, NULL, NULL);
  free(ops);
}

/* qsystem.c:464-489 */
void add_lin_term_list(qsystem sys, PetscScalar a, PetscInt num_ops, operator *ops) {
  sys->dm_equations = PETSC_TRUE;
  add_term(sys->time_indep, &sys->num_time_indep, sys->alloc_time_indep, LINDBLAD, a, num_ops, ops, NULL, NULL);
}

/* qsystem.c:491-521 */
void add_ham_term_time_dep(qsystem sys, PetscScalar a, void *ctx, PetscScalar (*f)(PetscReal, void *),
                           PetscInt num_ops, ...) {
  va_list ap;
  va_start(ap, num_ops);
  operator *ops = collect(num_ops, ap);
  va_end(ap);
  add_term(sys->time_dep, &sys->num_time_dep, sys->alloc_time_dep, TD_HAM, a, num_ops, ops, ctx, f);
  free(ops);
}

/* qsystem.c:523-552 */
void add_lin_term_time_dep(qsystem sys, PetscScalar a, void *ctx, PetscScalar (*f)(PetscReal, void *),
                           PetscInt num_ops, ...) {
  va_list ap;
  va_start(ap, num_ops);
  operator *ops = collect(num_ops, ap);
  va_end(ap);
  sys->dm_equations = PETSC_TRUE;
  add_term(sys->time_dep, &sys->num_time_dep, sys->alloc_time_dep, TD_LINDBLAD, a, num_ops, ops, ctx, f);
  free(ops);
}

void add_ham_term_time_dep_list(qsystem sys, PetscScalar a, void *ctx, quac_tdfunc_ri f, PetscInt num_ops,
                                operator *ops) {
  PetscInt before = sys->num_time_dep;
  add_term(sys->time_dep, &sys->num_time_dep, sys->alloc_time_dep, TD_HAM, a, num_ops, ops, ctx, NULL);
  if (sys->num_time_dep > before) sys->time_dep[before].time_dep_func_ri = f;
}

void add_lin_term_time_dep_list(qsystem sys, PetscScalar a, void *ctx, quac_tdfunc_ri f, PetscInt num_ops,
                                operator *ops) {
  PetscInt before = sys->num_time_dep;
  sys->dm_equations = PETSC_TRUE;
  add_term(sys->time_dep, &sys->num_time_dep, sys->alloc_time_dep, TD_LINDBLAD, a, num_ops, ops, ctx, NULL);
  if (sys->num_time_dep > before) sys->time_dep[before].time_dep_func_ri = f;
}

/* qsystem.c:557-591 */
void clear_mat_terms_sys(qsystem qsys) {
  for (PetscInt i = 0; i < qsys->num_time_dep; i++) {
    free(qsys->time_dep[i].ops);
    if (qsys->time_dep[i].mat_A) qh_mat_destroy(qsys->time_dep[i].mat_A);
    qsys->time_dep[i].mat_A = NULL;
  }
  for (PetscInt i = 0; i < qsys->num_time_indep; i++) free(qsys->time_indep[i].ops);
  if (qsys->mat_allocated) {
    qh_mat_destroy(qsys->mat_A);
    qsys->mat_A = NULL;
    qsys->mat_allocated = 0;
  }
  qsys->num_time_indep = 0;
  qsys->num_time_dep = 0;
}

/* ------------------------------------------------------------ assembly */
/* kind of a term under the dispatch of _add_ops_to_mat (kron.c:953-993) */
static int term_kind(qsystem s, mat_term_type t) {
  int lin = (t == LINDBLAD || t == TD_LINDBLAD);
  if (s->dm_equations) {
    if (s->mcwf_solver) return lin ? QH_LIN_MCWF : QH_HAM_ONLY;
    return lin ? QH_LIN_DM : QH_HAM_DM;
  }
  return lin ? -1 : QH_HAM_ONLY;
}

/* append one term's descriptor (MCWF Lindblad: reversed daggers then ops, kron.c:675-686) */
static int push_term_kind(qsystem s, qc_chain *c, qh_term **terms, int *nt, int *cap, mat_term *mt, int zero_a,
                          double are, double aim, int kind);
static int push_term(qsystem s, qc_chain *c, qh_term **terms, int *nt, int *cap, mat_term *mt, int zero_a,
                     double are, double aim) {
  return push_term_kind(s, c, terms, nt, cap, mt, zero_a, are, aim, term_kind(s, mt->my_term_type));
}

static int push_term_kind(qsystem s, qc_chain *c, qh_term **terms, int *nt, int *cap, mat_term *mt, int zero_a,
                          double are, double aim, int kind) {
  if (kind < 0) return quac_fail(0, "Wrong mat_term_type in construct_matrix - ham!"), -1;
  int op0 = c->nops;
  int n;
  if (kind == QH_LIN_MCWF) {
    PetscInt m = mt->num_ops;
    operator *tmp = (operator *)malloc(sizeof(operator) * (2 * m > 0 ? 2 * m : 1));
    for (PetscInt i = 0; i < m; i++) {
      tmp[m - i - 1] = qc_dag(mt->ops[i]);
      tmp[m + i] = mt->ops[i];
    }
    n = qc_chain_append(c, s->total_levels, 2 * m, tmp);
    free(tmp);
  } else {
    n = qc_chain_append(c, s->total_levels, mt->num_ops, mt->ops);
  }
  if (n < 0) return -1;
  if (*nt == *cap) {
    *cap = *cap ? 2 * *cap : 32;
    *terms = (qh_term *)realloc(*terms, sizeof(qh_term) * (size_t)*cap);
  }
  qh_term *t = &(*terms)[(*nt)++];
  t->kind = kind;
  t->nops = n;
  t->op0 = op0;
  t->zero_a = zero_a;
  t->a_re = are;
  t->a_im = aim;
  return 0;
}

static int assemble(qsystem s, int td_index, Mat *out) {
  qc_chain c;
  qc_chain_init(&c);
  qh_term *terms = NULL;
  int nt = 0, cap = 0, rc = 0;
  if (td_index >= 0) {
    rc = push_term(s, &c, &terms, &nt, &cap, &s->time_dep[td_index], 0, 1.0, 0.0);
  } else {
    for (PetscInt i = 0; i < s->num_time_indep && !rc; i++)
      rc = push_term(s, &c, &terms, &nt, &cap, &s->time_indep[i], 0, creal(s->time_indep[i].a),
                     cimag(s->time_indep[i].a));
    /* time-dependent terms seed the structure with a = 0 (qsystem.c:645-653) */
    for (PetscInt i = 0; i < s->num_time_dep && !rc; i++)
      rc = push_term(s, &c, &terms, &nt, &cap, &s->time_dep[i], 1, 0.0, 0.0);
  }
  if (!rc) {
    int flags = 0;
    if (td_index < 0 && (!s->dm_equations || s->mcwf_solver)) flags |= QH_ASM_SCALE_MINUS_I;
    if (quac_opt_int("quac_kron_conj_a", 0)) flags |= QH_ASM_CONJ_A;
    int r = qh_assemble(s->total_levels, s->dim, nt, terms, c.nops, c.ops, c.nlev, c.lev, flags, out);
    if (r) rc = quac_fail(9, "device assembly failed: %s", qh_last_error());
  }
  free(terms);
  qc_chain_free(&c);
  return rc;
}

/* The N x N effective Hamiltonian H - (i/2) sum a C^dag C of the time-independent
 * terms (the MCWF assembly of kron.c:638-700 without the -i scale), used by the
 * steady-state preconditioner. */
int quac_assemble_heff(qsystem s, Mat *out) {
  qc_chain c;
  qc_chain_init(&c);
  qh_term *terms = NULL;
  int nt = 0, cap = 0, rc = 0;
  for (PetscInt i = 0; i < s->num_time_indep && !rc; i++) {
    mat_term *mt = &s->time_indep[i];
    int lin = (mt->my_term_type == LINDBLAD || mt->my_term_type == TD_LINDBLAD);
    rc = push_term_kind(s, &c, &terms, &nt, &cap, mt, 0, creal(mt->a), cimag(mt->a), lin ? QH_LIN_MCWF : QH_HAM_ONLY);
  }
  if (!rc && nt == 0) rc = quac_fail(0, "no time-independent terms for the effective Hamiltonian");
  if (!rc) {
    int r = qh_assemble(s->total_levels, s->total_levels, nt, terms, c.nops, c.ops, c.nlev, c.lev, 0, out);
    if (r) rc = quac_fail(9, "device assembly of the effective Hamiltonian failed: %s", qh_last_error());
  }
  free(terms);
  qc_chain_free(&c);
  return rc;
}

/* qsystem.c:593-676 */
void construct_matrix(qsystem qsys) {
  qsys->hspace_frozen = 1;
  if (qsys->num_subsystems == 0) QUAC_FAIL(0, "You need to create operators before you construct a matrix!");
  if (qsys->num_time_dep + qsys->num_time_indep == 0)
    QUAC_FAIL(0, "You need to add some terms before you construct a matrix!");
  if (qsys->dm_equations) {
    qsys->dim = qsys->total_levels * qsys->total_levels;
    if (qsys->mcwf_solver) qsys->dim = qsys->total_levels;
  } else {
    if (qsys->mcwf_solver) QUAC_FAIL(9, "It does not make sense to use the mcwf_solver with no Lindblad terms.");
    qsys->dim = qsys->total_levels;
  }
  if (qsys->mat_allocated) {
    qh_mat_destroy(qsys->mat_A);
    qsys->mat_A = NULL;
    qsys->mat_allocated = 0;
  }
  qsys->my_num = qsys->dim;
  qsys->Istart = 0;
  qsys->Iend = qsys->dim;
  if (assemble(qsys, -1, &qsys->mat_A)) return;
  qsys->mat_allocated = 1;
  for (PetscInt i = 0; i < qsys->num_time_dep; i++) {
    if (qsys->time_dep[i].mat_A) qh_mat_destroy(qsys->time_dep[i].mat_A);
    qsys->time_dep[i].mat_A = NULL;
    if (assemble(qsys, (int)i, &qsys->time_dep[i].mat_A)) return;
  }
}

/* qsystem.c:836-863: the N x N matrix a * ops[0]*ops[1]*... of the WF space */
void construct_op_matrix_wf_list(qsystem qsys, PetscScalar a, Mat *mat, PetscInt num_ops, operator *ops) {
  if (qsys->num_subsystems == 0) QUAC_FAIL(0, "You need to create operators before you construct a matrix!");
  qc_chain c;
  qc_chain_init(&c);
  int n = qc_chain_append(&c, qsys->total_levels, num_ops, ops);
  if (n < 0) {
    qc_chain_free(&c);
    return;
  }
  qh_term t = {QH_HAM_ONLY, n, 0, 0, creal(a), cimag(a)};
  int r = qh_assemble(qsys->total_levels, qsys->total_levels, 1, &t, c.nops, c.ops, c.nlev, c.lev, 0, mat);
  qc_chain_free(&c);
  if (r) QUAC_FAIL(9, "device assembly failed: %s", qh_last_error());
}

/* ------------------------------------------------------------ stepping */
struct quac_ts {
  qsystem sys;
  PetscReal t;
  PetscInt step;
};

void TSGetTime(TS ts, PetscReal *t) { *t = ts->t; }
void TSGetStepNumber(TS ts, PetscInt *s) { *s = ts->step; }

static PetscInt g_last_steps = 0;
static PetscReal g_last_t = 0.0;
void quac_ts_stats(PetscInt *steps, PetscReal *t_final) {
  if (steps) *steps = g_last_steps;
  if (t_final) *t_final = g_last_t;
}

/* qsystem.c:875-879 */
void set_ts_monitor_sys(qsystem sys, PetscErrorCode (*monitor)(TS, PetscInt, PetscReal, Vec, void *), void *ctx) {
  sys->ts_monitor = monitor;
  sys->ts_ctx = ctx;
}

static int call_monitor(qsystem s, struct quac_ts *ts, qvec x) {
  if (!s->ts_monitor) return 0;
  ts->sys = s;
  return s->ts_monitor(ts, ts->step, ts->t, x->data, s->ts_ctx);
}

/* monitor of the MCWF ensemble: called once per ensemble step with the
 * ensemble's storage as the Vec (all trajectories, trajectory i at [i*N, (i+1)*N)) */
static int mcwf_monitor(qsystem s, PetscInt step, PetscReal t, qvec x) {
  struct quac_ts ts = {s, t, step};
  ts.t = t;
  ts.step = step;
  return s->ts_monitor(&ts, step, t, x->ens_spawned ? x->ens_datas[0] : x->data, s->ts_ctx);
}

/* s_k(t) = a_k f_k(t) of every time-dependent term, times -i for a WF or MCWF
 * generator (qsystem.c:1300-1330), as re/im pairs; mats gets the A_k */
static void td_coefs(qsystem s, PetscReal t, double *sc, Mat *mats) {
  for (PetscInt k = 0; k < s->num_time_dep; k++) {
    PetscScalar f = 1.0;
    if (s->time_dep[k].time_dep_func) {
      f = s->time_dep[k].time_dep_func(t, s->time_dep[k].ctx);
    } else if (s->time_dep[k].time_dep_func_ri) {
      PetscReal fr = 0.0, fi = 0.0;
      s->time_dep[k].time_dep_func_ri(t, s->time_dep[k].ctx, &fr, &fi);
      f = fr + fi * PETSC_i;
    }
    PetscScalar v = s->time_dep[k].a * f;
    if (!s->dm_equations || s->mcwf_solver) v = -PETSC_i * v;
    sc[2 * k] = creal(v);
    sc[2 * k + 1] = cimag(v);
    if (mats) mats[k] = s->time_dep[k].mat_A;
  }
}

/* rho <- rho / Tr(rho) (FP64 trace of the FP32 state) */
static int renorm_trace(Vec x, PetscInt tot) {
  double re = 0.0, im = 0.0;
  int rc = qh_trace(x, tot, &re, &im);
  if (!rc && re != 0.0) rc = qh_vec_scale(x, 1.0 / re, 0.0);
  return rc;
}

/* one classic RK4 step of length h at time t with the time-dependent terms: the
 * envelopes at the four stage times (each evaluated once per stage, as the
 * reference's RHS callback is), then four fused stage launches */
static int rk4_td_step(qsystem s, Vec x, PetscReal t, PetscReal h) {
  const int n = (int)s->num_time_dep;
  double sc[4 * 2 * NUM_INIT_ALLOC];
  Mat mats[NUM_INIT_ALLOC];
  const PetscReal ts[4] = {t, t + 0.5 * h, t + 0.5 * h, t + h};
  for (int j = 0; j < 4; j++) td_coefs(s, ts[j], sc + 2 * (size_t)j * n, mats);
  return qh_rk4_td_step(s->mat_A, n, mats, sc, x, h);
}

/* nfull RK4 steps of length h from time t with the time-dependent terms: the
 * envelopes at the four stage times of every step evaluated here (in step
 * order, once per stage), the steps run in chunks of up to 256 by
 * qh_rk4_td_steps (one workgroup for a generator of at most 800 rows) */
static int rk4_td_run(qsystem s, Vec x, PetscReal t, PetscReal h, PetscInt nfull) {
  const int n = (int)s->num_time_dep;
  const PetscInt chunk = nfull < 256 ? nfull : 256;
  double *sc = malloc(sizeof(double) * 8 * (size_t)n * (size_t)chunk + sizeof(double));
  Mat mats[NUM_INIT_ALLOC];
  if (!sc) return 1;
  int rc = 0;
  for (PetscInt j0 = 0; j0 < nfull && !rc; j0 += chunk) {
    const PetscInt m = nfull - j0 < chunk ? nfull - j0 : chunk;
    for (PetscInt j = 0; j < m; j++) {
      const PetscReal tj = t + (double)(j0 + j) * h;
      const PetscReal ts[4] = {tj, tj + 0.5 * h, tj + 0.5 * h, tj + h};
      for (int q = 0; q < 4; q++) td_coefs(s, ts[q], sc + 2 * (size_t)n * (4 * (size_t)j + q), mats);
    }
    rc = qh_rk4_td_steps(s->mat_A, n, mats, sc, m, x, h);
  }
  free(sc);
  return rc;
}

typedef struct {
  Vec k[4], y, xn;
} work_t;

static int work_alloc(work_t *w, PetscInt n) {
  memset(w, 0, sizeof(*w));
  for (int i = 0; i < 4; i++)
    if (qh_vec_create(n, QH_FP64, &w->k[i])) return -1;
  if (qh_vec_create(n, QH_FP64, &w->y) || qh_vec_create(n, QH_FP64, &w->xn)) return -1;
  return 0;
}
static void work_free(work_t *w) {
  for (int i = 0; i < 4; i++) qh_vec_destroy(w->k[i]);
  qh_vec_destroy(w->y);
  qh_vec_destroy(w->xn);
}

/*
 * Bogacki-Shampine 3(2) step with FSAL (PETSc TSRK3BS tableau); returns the
 * weighted RMS error estimate in *err and the candidate in w->xn.
 */
static int rk3bs_try(qsystem s, Vec x, PetscReal t, PetscReal h, work_t *w, int have_k1, double atol, double rtol,
                     double *err) {
  int rc = 0;
  int td = s->num_time_dep > 0;
  if (!td) /* time-independent RHS: the fused device step */
    return qh_rk3bs_step(s->mat_A, x, w->xn, w->k[0], w->k[1], w->k[2], w->k[3], w->y, h, atol, rtol, have_k1, err);
  /* the envelopes at the four stage times, each evaluated once per stage as the
   * reference's RHS callback is (the first not at all under FSAL) */
  const int n = (int)s->num_time_dep;
  double sc[4 * 2 * NUM_INIT_ALLOC];
  Mat mats[NUM_INIT_ALLOC];
  const PetscReal ts[4] = {t, t + 0.5 * h, t + 0.75 * h, t + h};
  for (int j = 0; j < 4; j++) {
    if (j == 0 && have_k1) {
      memset(sc, 0, sizeof(double) * 2 * (size_t)n);
      continue;
    }
    td_coefs(s, ts[j], sc + 2 * (size_t)j * n, mats);
  }
  rc = qh_rk3bs_td_step(s->mat_A, n, mats, sc, x, w->xn, w->k[0], w->k[1], w->k[2], w->k[3], w->y, h, atol, rtol,
                        have_k1, err);
  return rc;
}

/* qsystem.c:895-1098 */
void time_step_sys(qsystem qsys, qvec x, PetscReal init_time, PetscReal time_max, PetscReal dt, PetscInt steps_max) {
  qsys->time_step_called = PETSC_TRUE;
  qsys->solution_qvec = x;
  if (qsys->mat_allocated != 1) QUAC_FAIL(0, "Matrix does not seem to be constructed! You need to call construct_matrix!");
  if (!x || !x->data) QUAC_FAIL(0, "time_step_sys: state not created");
  if (x->n != qsys->dim) QUAC_FAIL(0, "time_step_sys: state size %lld does not match the system dimension %lld",
                                   (long long)x->n, (long long)qsys->dim);
  if (dt <= 0) QUAC_FAIL(0, "time_step_sys: dt must be positive");
  if (qsys->mcwf_solver && qsys->num_time_dep > 0)
    QUAC_FAIL(9, "time_step_sys: time-dependent terms with use_mcwf_solver are not supported (the trajectories "
                 "step with the time-independent effective Hamiltonian on the device)");
  if (qsys->mcwf_solver) {
    /* trajectories of the ensemble (qsystem.c:1024-1079); a circuit's gates
     * are events (qsystem.c:1006-1019, quantum_circuits.c:17-120): the ensemble
     * runs to each gate time, where every trajectory takes the gate, and then
     * on (steps of dt again from the event) */
    const PetscReal eps = 1e-12 * (fabs(time_max) + fabs(dt) + 1.0);
    PetscReal t = init_time;
    PetscInt total = 0;
    int first = 1;
    for (;;) {
      PetscReal t_ev = 0.0, tend = time_max;
      const int have_ev = qsys->num_circuits > 0 && qc_next_event_time(qsys, &t_ev) && t_ev > t + eps &&
                          t_ev < time_max + eps;
      if (have_ev && t_ev < tend) tend = t_ev;
      PetscInt nsteps = 0;
      if (quac_mcwf_time_step(qsys, x, t, tend, dt, steps_max - total, qsys->ts_monitor ? mcwf_monitor : NULL,
                              &nsteps, total, first))
        QUAC_FAIL(9, "time_step_sys (MCWF) failed");
      first = 0;
      total += nsteps;
      t = qsys->current_time;
      if (!(have_ev && fabs(t - t_ev) <= 1e-9)) break;
      t = t_ev;
      if (qc_apply_events(qsys, t, x)) return;
      if (t >= time_max - eps || total >= steps_max) break;
    }
    g_last_steps = total;
    g_last_t = qsys->current_time;
    return;
  }

  /* integrator: PETSc default here is TSRK3BS adaptive (qsystem.c:1003-1005);
   * -ts_rk_type 4 -ts_adapt_type none selects fixed-step classic RK4 */
  const char *rk = quac_opt_str("ts_rk_type", "3bs");
  const char *adapt = quac_opt_str("ts_adapt_type", "basic");
  const int rk4 = !strcmp(rk, "4");
  const int fixed = rk4 || !strcmp(adapt, "none");
  const int fp32 = !strcmp(quac_opt_str("quac_precision", "fp64"), "fp32");
  const int renorm = quac_opt_int("quac_renorm", 0);
  const int td = qsys->num_time_dep > 0;
  const int events = qsys->num_circuits > 0;

  struct quac_ts ts = {qsys, init_time, 0};
  PetscReal t = init_time;
  PetscInt step = 0;
  const PetscReal eps = 1e-12 * (fabs(time_max) + fabs(dt) + 1.0);
  if (qsys->ts_monitor && call_monitor(qsys, &ts, x)) QUAC_FAIL(9, "monitor returned an error");

  if (rk4 && fixed) {
    /* fused device RK4 (FP64 drop-in path, or the FP32 bulk path); with
     * time-dependent terms one fused-RHS step at a time */
    Vec work = x->data;
    Vec w32 = NULL;
    if (fp32) {
      QUAC_HIPCHK(qh_vec_create(x->n, QH_FP32, &w32));
      QUAC_HIPCHK(qh_vec_copy(x->data, w32));
      work = w32;
    }
    const int renorm_ok = fp32 && renorm > 0 && x->my_type == DENSITY_MATRIX;
    while (t < time_max - eps && step < steps_max) {
      PetscReal t_ev = 0.0;
      int have_ev = events && qc_next_event_time(qsys, &t_ev) && t_ev > t + eps;
      PetscReal h = dt;
      PetscInt nfull = 1;
      if (!qsys->ts_monitor) {
        /* as many full steps as fit before tmax, the next event and steps_max */
        PetscReal lim = time_max;
        if (have_ev && t_ev < lim) lim = t_ev;
        double q = (lim - t) / dt;
        nfull = (PetscInt)floor(q + 1e-9);
        if (nfull > steps_max - step) nfull = steps_max - step;
        if (nfull < 1) {
          nfull = 1;
          h = lim - t;
        }
      } else {
        if (t + h > time_max - eps) h = time_max - t;
        if (have_ev && t + h > t_ev - eps) h = t_ev - t;
      }
      if (td) {
        const PetscReal hs = (nfull == 1 && h != dt) ? h : dt;
        int bad = 0;
        if (!renorm_ok) bad = rk4_td_run(qsys, work, t, hs, nfull);
        for (PetscInt j = 0; j < nfull && !bad && renorm_ok; j++) {
          bad = rk4_td_step(qsys, work, t + (double)j * hs, hs);
          /* FP32: trace renormalisation every quac_renorm steps */
          if (!bad && renorm_ok && (step + j + 1) % renorm == 0) bad = renorm_trace(work, x->total_levels);
        }
        if (bad) break;
        t += (double)nfull * hs;
        if (have_ev && fabs(t - t_ev) < 1e-9) t = t_ev;
      } else if (nfull == 1 && h != dt) {
        if (qh_rk4(qsys->mat_A, work, h, 1, renorm_ok ? x->total_levels : 0, 0)) break;
        t += h;
      } else {
        if (qh_rk4(qsys->mat_A, work, dt, nfull, renorm_ok ? x->total_levels : 0, renorm_ok ? renorm : 0)) break;
        t += (double)nfull * dt;
        if (have_ev && fabs(t - t_ev) < 1e-9) t = t_ev;
      }
      step += nfull;
      if (events && have_ev && fabs(t - t_ev) <= 1e-9) {
        /* the FP32 path applies the gates to its hi+lo state directly (the gate
         * kernel computes in FP64 and splits the result again) */
        Vec keep = x->data;
        x->data = work;
        const int bad = qc_apply_events(qsys, t, x);
        x->data = keep;
        if (bad) break;
      }
      if (qsys->ts_monitor) {
        if (fp32) QUAC_HIPCHK(qh_vec_copy(w32, x->data));
        ts.t = t;
        ts.step = step;
        if (call_monitor(qsys, &ts, x)) QUAC_FAIL(9, "monitor returned an error");
      }
    }
    if (fp32) {
      QUAC_HIPCHK(qh_vec_copy(w32, x->data));
      qh_vec_destroy(w32);
    }
  } else {
    /* host-driven RK3BS on device vectors, adaptive (atol = rtol = 5e-7, dt
     * limits [0.01, 1]: qsystem.c:902,995) or at a fixed step */
    work_t w;
    if (work_alloc(&w, x->n)) {
      work_free(&w);
      QUAC_FAIL(9, "time_step_sys: out of device memory");
    }
    const double atol = quac_opt_double("ts_atol", 5e-7), rtol = quac_opt_double("ts_rtol", 5e-7);
    const double hmin = quac_opt_double("ts_adapt_dt_min", 0.01), hmax = quac_opt_double("ts_adapt_dt_max", 1.0);
    PetscReal h = dt;
    int have_k1 = 0, rc = 0;
    int64_t vn = 0;
    int vprec = 0;
    qh_vec_size(x->data, &vn, &vprec);
    if (!td && vprec == QH_FP64 && qsys->dim <= 640 && quac_opt_int("quac_small_path", 1)) {
      /* a small generator: the integration in one workgroup, controller
       * included (same formulas as the loop below); up to the next event, and
       * with a monitor one accepted step per launch (the monitor is called after
       * accepted steps, as PETSc's TSMonitor is) */
      PetscReal hcur = dt;
      while (t < time_max - eps && step < steps_max && !rc) {
        PetscReal t_ev = 0.0, tend = time_max;
        const int have_ev = events && qc_next_event_time(qsys, &t_ev) && t_ev > t + eps;
        if (have_ev && t_ev < tend) tend = t_ev;
        double tf = t, hl = hcur;
        int64_t ns = 0;
        /* at most 65 536 accepted steps per launch, so that no launch runs for long */
        const int64_t room = steps_max - step < 65536 ? steps_max - step : 65536;
        rc = qh_rk3bs_small(qsys->mat_A, x->data, t, tend, hcur, qsys->ts_monitor ? 1 : room, atol, rtol, hmin, hmax,
                            !fixed, &tf, &ns, &hl);
        if (rc) break;
        t = tf;
        step += ns;
        hcur = fixed ? dt : hl;
        const int at_ev = have_ev && fabs(t - t_ev) <= 1e-9;
        if (at_ev) {
          t = t_ev;
          if (qc_apply_events(qsys, t, x)) rc = 1;
        }
        if (!rc && qsys->ts_monitor) {
          ts.t = t;
          ts.step = step;
          if (call_monitor(qsys, &ts, x)) rc = 9;
        }
        if (ns == 0 && !at_ev) break;
      }
    }
    while (t < time_max - eps && step < steps_max && !rc) {
      PetscReal t_ev = 0.0;
      int have_ev = events && qc_next_event_time(qsys, &t_ev) && t_ev > t + eps;
      PetscReal hh = h;
      if (t + hh > time_max) hh = time_max - t;
      if (have_ev && t + hh > t_ev) hh = t_ev - t;
      if (fixed) {
        double err;
        rc = rk3bs_try(qsys, x->data, t, hh, &w, 0, atol, rtol, &err);
        if (!rc) rc = qh_vec_copy(w.xn, x->data);
        t += hh;
        step++;
      } else {
        double err = 0.0;
        rc = rk3bs_try(qsys, x->data, t, hh, &w, have_k1, atol, rtol, &err);
        if (rc) break;
        if (!(err <= 1.7976931348623157e308)) {
          /* NaN or infinite: no step size can be accepted */
          work_free(&w);
          QUAC_FAIL(9, "time_step_sys: non-finite error estimate at t = %g (the state is not finite)", t);
        }
        double fac = 0.9 * pow(err > 1e-300 ? err : 1e-300, -1.0 / 3.0);
        if (fac < 0.1) fac = 0.1;
        if (fac > 10.0) fac = 10.0;
        double hnew = hh * fac;
        if (hnew < hmin) hnew = hmin;
        if (hnew > hmax) hnew = hmax;
        if (err <= 1.0 || hh <= hmin * (1 + 1e-12)) {
          rc = qh_vec_copy(w.xn, x->data);
          Vec tmp = w.k[0];
          w.k[0] = w.k[3];
          w.k[3] = tmp;
          t += hh;
          step++;
          if (hh < h && have_ev && fabs(t - t_ev) <= 1e-9) hnew = h; /* shortened for the event */
        }
        /* FSAL: k[0] now holds f(t, x) for the current t and x (accepted: the
         * last stage f(t + h, x_{n+1}), its time-dependent coefficients taken at
         * t + h; rejected: unchanged) */
        have_k1 = 1;
        h = hnew;
      }
      if (!rc && events && have_ev && fabs(t - t_ev) <= 1e-9) {
        if (qc_apply_events(qsys, t, x)) rc = 1;
        have_k1 = 0;
      }
      if (!rc && qsys->ts_monitor) {
        ts.t = t;
        ts.step = step;
        if (call_monitor(qsys, &ts, x)) rc = 9;
      }
    }
    work_free(&w);
    if (rc) QUAC_FAIL(9, "time_step_sys failed: %s", qh_last_error());
  }
  qh_synchronize();
  g_last_steps = step;
  g_last_t = t;
  if (fabs(t - time_max) > 1e-8) PetscPrintf(0, "ERROR! Did not reach final time! Reason: %d\n", -1);
}

/* ------------------------------------------------------------ batched sweeps */
/* The batch of the last time_step_sys_batch call, kept for the next call with
 * the same matrices (by serial number) and precision. */
static struct {
  qh_batch *b;
  PetscInt n;
  uint64_t *uid;
  int prec;
} g_batch = {NULL, 0, NULL, 0};

void time_step_sys_batch_release(void) {
  if (g_batch.b) qh_batch_destroy(g_batch.b);
  free(g_batch.uid);
  g_batch.b = NULL;
  g_batch.uid = NULL;
  g_batch.n = 0;
}

/* n independent density-matrix systems advanced together by fixed-step classic
 * RK4 from init_time to time_max (steps of dt, the last one shortened to land
 * on time_max, at most steps_max): one block-diagonal device matrix, one fused
 * stage kernel per RK stage over every instance, hipGraph replay.  The systems
 * must be constructed density-matrix systems without time-dependent terms,
 * circuits or the MCWF solver; x[i] is advanced in place.  Precision:
 * -quac_batch_precision {fp32 (default), fp64}; FP32 rescales each instance to
 * unit trace every -quac_renorm steps (0: never).  Monitors are not called. */
void time_step_sys_batch(PetscInt n, qsystem *sys, qvec *x, PetscReal init_time, PetscReal time_max, PetscReal dt,
                         PetscInt steps_max) {
  if (n < 1 || !sys || !x) QUAC_FAIL(0, "time_step_sys_batch: no systems");
  if (dt <= 0) QUAC_FAIL(0, "time_step_sys_batch: dt must be positive");
  for (PetscInt i = 0; i < n; i++) {
    qsystem s = sys[i];
    if (!s || s->mat_allocated != 1)
      QUAC_FAIL(0, "time_step_sys_batch: system %lld is not constructed (call construct_matrix)", (long long)i);
    if (!s->dm_equations || s->mcwf_solver)
      QUAC_FAIL(0, "time_step_sys_batch: system %lld is not a density-matrix system", (long long)i);
    if (s->num_time_dep > 0 || s->num_circuits > 0)
      QUAC_FAIL(0, "time_step_sys_batch: system %lld has time-dependent terms or circuits", (long long)i);
    if (!x[i] || !x[i]->data || x[i]->n != s->dim)
      QUAC_FAIL(0, "time_step_sys_batch: state %lld does not match its system", (long long)i);
  }
  const int prec = strcmp(quac_opt_str("quac_batch_precision", "fp32"), "fp64") ? QH_FP32 : QH_FP64;
  const int renorm = prec == QH_FP32 ? quac_opt_int("quac_renorm", 0) : 0;
  int same = g_batch.b && g_batch.n == n && g_batch.prec == prec;
  for (PetscInt i = 0; same && i < n; i++) same = g_batch.uid[i] == qh_mat_uid(sys[i]->mat_A);
  if (!same) {
    time_step_sys_batch_release();
    qh_mat **mats = (qh_mat **)malloc(sizeof(qh_mat *) * (size_t)n);
    int64_t *tot = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    g_batch.uid = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    if (!mats || !tot || !g_batch.uid) {
      free(mats);
      free(tot);
      time_step_sys_batch_release();
      QUAC_FAIL(9, "time_step_sys_batch: out of host memory");
    }
    for (PetscInt i = 0; i < n; i++) {
      mats[i] = sys[i]->mat_A;
      tot[i] = sys[i]->total_levels;
      g_batch.uid[i] = qh_mat_uid(sys[i]->mat_A);
    }
    const int rc = qh_batch_create((int)n, mats, tot, prec, &g_batch.b);
    free(mats);
    free(tot);
    if (rc) {
      g_batch.b = NULL;
      time_step_sys_batch_release();
      QUAC_FAIL(9, "time_step_sys_batch: %s", qh_last_error());
    }
    g_batch.n = n;
    g_batch.prec = prec;
  }
  for (PetscInt i = 0; i < n; i++) QUAC_HIPCHK(qh_batch_set_state_vec(g_batch.b, (int)i, x[i]->data));
  const PetscReal eps = 1e-12 * (fabs(time_max) + fabs(dt) + 1.0);
  PetscReal t = init_time;
  PetscInt steps = 0;
  if (time_max > t + eps && steps_max > 0) {
    PetscInt nfull = (PetscInt)floor((time_max - t) / dt + 1e-9);
    if (nfull > steps_max) nfull = steps_max;
    if (nfull > 0) {
      QUAC_HIPCHK(qh_batch_rk4(g_batch.b, dt, nfull, renorm));
      t += (double)nfull * dt;
      steps = nfull;
    }
    if (steps < steps_max && time_max - t > eps) {
      QUAC_HIPCHK(qh_batch_rk4(g_batch.b, time_max - t, 1, 0));
      t = time_max;
      steps++;
    }
  }
  for (PetscInt i = 0; i < n; i++) {
    QUAC_HIPCHK(qh_batch_get_state_vec(g_batch.b, (int)i, x[i]->data));
    sys[i]->time_step_called = PETSC_TRUE;
    sys[i]->current_time = t;
  }
  QUAC_HIPCHK(qh_synchronize());
  g_last_steps = steps;
  g_last_t = t;
}
