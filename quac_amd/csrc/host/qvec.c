/* qvec.c -- state containers and observables of the QuaC-compatible host layer
 * (src/qvec_utilities.c; line numbers cited per function).  The data live in
 * HBM; observables are reduced in FP64 on the device. */
#include <math.h>
#include <string.h>

#include "quac_internal.h"

static qvec new_qvec(qvec_type type, PetscInt n, PetscInt total_levels) {
  qvec q = (qvec)calloc(1, sizeof(struct qvec));
  q->my_type = type;
  q->n = n;
  q->Istart = 0;
  q->Iend = n;
  q->total_levels = total_levels;
  q->n_ensemble = -1;
  q->ens_i = -1;
  q->ens_spawned = PETSC_FALSE;
  if (qh_vec_create(n, QH_FP64, &q->data)) {
    free(q);
    quac_fail(9, "cannot allocate a state of %lld entries: %s", (long long)n, qh_last_error());
    return NULL;
  }
  return q;
}

/* qvec_utilities.c:633-648 */
void create_qvec_sys(qsystem qsys, qvec *new_qvec) {
  if (qsys->dm_equations == PETSC_TRUE && qsys->mcwf_solver == PETSC_FALSE) create_dm_sys(qsys, new_qvec);
  else if (qsys->mcwf_solver == PETSC_TRUE) create_wf_ensemble_sys(qsys, new_qvec);
  else create_wf_sys(qsys, new_qvec);
}

/* qvec_utilities.c:765-817: the initial condition lives in data; the
 * trajectories are spawned from it by the first time_step_sys, into one device
 * block of n_ensemble x N (ens_datas[0]) */
void create_wf_ensemble_sys(qsystem qsys, qvec *new_wf) {
  if (quac_ensure_device()) return;
  if (qsys->num_subsystems == 0) QUAC_FAIL(0, "You need to create operators before you construct a wavefunction!");
  if (qsys->dm_equations == PETSC_TRUE && qsys->mcwf_solver == PETSC_FALSE)
    QUAC_FAIL(0, "Must use density matrices if Lindblad terms are used without using mcwf_solver.");
  qsys->dim = qsys->total_levels;
  qsys->hspace_frozen = 1;
  qvec q = new_qvec(WF_ENSEMBLE, qsys->dim, qsys->total_levels);
  if (!q) return;
  q->n_ops = qsys->num_subsystems;
  q->ndims_hspace = qsys->num_subsystems;
  q->hspace_dims = (PetscInt *)malloc(qsys->num_subsystems * sizeof(PetscInt));
  for (PetscInt i = 0; i < qsys->num_subsystems; i++) q->hspace_dims[i] = qsys->subsystem_list[i]->my_levels;
  q->n_ensemble = qsys->num_local_trajs;
  q->ens_spawned = PETSC_FALSE;
  *new_wf = q;
}

/* qvec_utilities.c:654-705 */
void create_dm_sys(qsystem qsys, qvec *new_dm) {
  if (quac_ensure_device()) return;
  if (qsys->num_subsystems == 0) QUAC_FAIL(0, "You need to create operators before you construct a density matrix!");
  qsys->dm_equations = PETSC_TRUE;
  qsys->dim = qsys->total_levels * qsys->total_levels;
  qsys->hspace_frozen = 1;
  qvec q = new_qvec(DENSITY_MATRIX, qsys->dim, qsys->total_levels);
  if (!q) return;
  q->n_ops = qsys->num_subsystems;
  q->ndims_hspace = 2 * qsys->num_subsystems;
  q->hspace_dims = (PetscInt *)malloc(2 * qsys->num_subsystems * sizeof(PetscInt));
  for (PetscInt i = 0; i < qsys->num_subsystems; i++) {
    q->hspace_dims[i] = qsys->subsystem_list[i]->my_levels;
    q->hspace_dims[i + qsys->num_subsystems] = qsys->subsystem_list[i]->my_levels;
  }
  *new_dm = q;
}

/* qvec_utilities.c:710-760 */
void create_wf_sys(qsystem qsys, qvec *new_wf) {
  if (quac_ensure_device()) return;
  if (qsys->num_subsystems == 0) QUAC_FAIL(0, "You need to create operators before you construct a wavefunction!");
  if (qsys->dm_equations == PETSC_TRUE && qsys->mcwf_solver == PETSC_FALSE)
    QUAC_FAIL(0, "Must use density matrices if Lindblad terms are used without using mcwf_solver.");
  qsys->dim = qsys->total_levels;
  qsys->hspace_frozen = 1;
  qvec q = new_qvec(WAVEFUNCTION, qsys->dim, qsys->total_levels);
  if (!q) return;
  q->n_ops = qsys->num_subsystems;
  q->ndims_hspace = qsys->num_subsystems;
  q->hspace_dims = (PetscInt *)malloc(qsys->num_subsystems * sizeof(PetscInt));
  for (PetscInt i = 0; i < qsys->num_subsystems; i++) q->hspace_dims[i] = qsys->subsystem_list[i]->my_levels;
  *new_wf = q;
}

void create_arb_qvec(qvec *out, PetscInt n, qvec_type type) {
  if (quac_ensure_device()) return;
  PetscInt tl = n;
  if (type == DENSITY_MATRIX) {
    tl = (PetscInt)llround(sqrt((double)n));
    if (tl * tl != n) QUAC_FAIL(9, "n for DENSITY_MATRIX must be perfect square!");
  }
  qvec q = new_qvec(type, n, tl);
  if (!q) return;
  q->n_ops = 1;
  q->ndims_hspace = type == DENSITY_MATRIX ? 2 : 1;
  q->hspace_dims = (PetscInt *)malloc(2 * sizeof(PetscInt));
  q->hspace_dims[0] = q->hspace_dims[1] = tl;
  *out = q;
}

/* qvec_utilities.c:523-570: a zero state with the given subsystem dimensions
 * (a density matrix lists every dimension twice: ndims = 2 x subsystems) */
void create_arb_qvec_dims(qvec *out, PetscInt ndims, PetscInt *dims, qvec_type type) {
  if (quac_ensure_device()) return;
  if (ndims < 1 || (type == DENSITY_MATRIX && ndims % 2)) QUAC_FAIL(9, "create_arb_qvec_dims: bad ndims %lld", (long long)ndims);
  if (type != WAVEFUNCTION && type != DENSITY_MATRIX) QUAC_FAIL(9, "ERROR! Type not understood in create_arb_qvec_dims.");
  PetscInt nstates = 1;
  for (PetscInt i = 0; i < ndims; i++) nstates *= dims[i];
  const PetscInt tl = type == DENSITY_MATRIX ? (PetscInt)llround(sqrt((double)nstates)) : nstates;
  qvec q = new_qvec(type, nstates, tl);
  if (!q) return;
  q->n_ops = type == DENSITY_MATRIX ? ndims / 2 : ndims;
  q->ndims_hspace = ndims;
  q->hspace_dims = (PetscInt *)malloc(sizeof(PetscInt) * (size_t)ndims);
  for (PetscInt i = 0; i < ndims; i++) q->hspace_dims[i] = dims[i];
  *out = q;
}

/* qvec_utilities.c:1076-1090 */
void destroy_qvec(qvec *state) {
  if (!state || !*state) return;
  free((*state)->hspace_dims);
  if ((*state)->ens_datas) {
    qh_vec_destroy((*state)->ens_datas[0]);
    free((*state)->ens_datas);
  }
  qh_vec_destroy((*state)->data);
  free(*state);
  *state = NULL;
}

/* qvec_utilities.c:879-891 */
void add_to_qvec_loc(qvec state, PetscScalar alpha, PetscInt loc) {
  if (loc > state->n || loc < 0) QUAC_FAIL(0, "Adding to a qvec location that is out of range!");
  if (loc == state->n) QUAC_FAIL(0, "Adding to a qvec location that is out of range!");
  QUAC_HIPCHK(qh_vec_add_entry(state->data, loc, creal(alpha), cimag(alpha)));
}

/* qvec_utilities.c:919-965: DM takes (i, j) with loc = sqrt(n)*i + j */
void add_to_qvec(qvec state, PetscScalar alpha, ...) {
  va_list ap;
  va_start(ap, alpha);
  PetscInt loc = va_arg(ap, PetscInt);
  if (state->my_type == DENSITY_MATRIX) {
    PetscInt loc2 = va_arg(ap, PetscInt);
    loc = state->total_levels * loc + loc2;
  }
  va_end(ap);
  add_to_qvec_loc(state, alpha, loc);
}

/* qvec_utilities.c:996-1033 (keeps the reference's n_before-weighted location) */
void get_qvec_loc_fock_op_list(qvec state, PetscInt *loc, PetscInt num_ops, operator *ops, PetscInt *fock) {
  *loc = 0;
  for (PetscInt i = 0; i < num_ops; i++) {
    if (fock[i] >= ops[i]->my_levels || fock[i] < 0)
      QUAC_FAIL(0, "The Fock State must be less than the number of levels and greater than 0!");
    *loc += fock[i] * ops[i]->n_before;
  }
  if (state->my_type == DENSITY_MATRIX) *loc = *loc * state->total_levels + *loc;
}

void add_to_qvec_fock_op_list(PetscScalar alpha, qvec state, PetscInt num_ops, operator *ops, PetscInt *fock) {
  PetscInt loc = 0;
  quac_clear_error();
  get_qvec_loc_fock_op_list(state, &loc, num_ops, ops, fock);
  if (quac_last_error()[0]) return;
  add_to_qvec_loc(state, alpha, loc);
}

/* qvec_utilities.c:817-842 */
void add_to_qvec_fock_op(PetscScalar alpha, qvec state, PetscInt num_ops, ...) {
  va_list ap;
  operator *ops = (operator *)malloc(sizeof(operator) * (num_ops > 0 ? num_ops : 1));
  PetscInt *fock = (PetscInt *)malloc(sizeof(PetscInt) * (num_ops > 0 ? num_ops : 1));
  va_start(ap, num_ops);
  for (PetscInt i = 0; i < num_ops; i++) {
    ops[i] = va_arg(ap, operator);
    fock[i] = va_arg(ap, PetscInt);
  }
  va_end(ap);
  add_to_qvec_fock_op_list(alpha, state, num_ops, ops, fock);
  free(ops);
  free(fock);
}

/* qvec_utilities.c:1047-1064: device writes are already in place; order them */
void assemble_qvec(qvec state) { (void)state; }

/* qvec_utilities.c:2540-2566 */
void set_qvec_from_init_excited_op(qsystem qsys, qvec state) {
  PetscInt init = 0;
  for (PetscInt i = 0; i < qsys->num_subsystems; i++) {
    PetscInt n_after = qsys->total_levels / (qsys->subsystem_list[i]->my_levels * qsys->subsystem_list[i]->n_before);
    init += qsys->subsystem_list[i]->initial_exc * n_after;
  }
  if (qsys->dm_equations == PETSC_TRUE) init = qsys->total_levels * init + init;
  add_to_qvec_loc(state, 1.0, init);
}

/* qsystem.c:1340-1349 */
void set_init_excited_op(operator op, PetscInt initial_exc) {
  if (initial_exc >= op->my_levels && op->my_op_type != VEC)
    QUAC_FAIL(0, "The initial excitation level must be less than the number of levels and can't be used for VEC_OP!");
  op->initial_exc = initial_exc;
}

/* qvec_utilities.c:1176-1195, 1197-1281 */
void get_expectation_value_qvec_list(qvec state, PetscScalar *trace_val, PetscInt num_ops, operator *ops) {
  qc_chain c;
  qc_chain_init(&c);
  if (qc_chain_append(&c, state->total_levels, num_ops, ops) < 0) {
    qc_chain_free(&c);
    return;
  }
  double re = 0.0, im = 0.0;
  int rc;
  if (state->my_type == DENSITY_MATRIX)
    rc = qh_expect_dm(state->data, state->total_levels, c.nops, c.ops, c.nlev, c.lev, &re, &im);
  else if (state->my_type == WF_ENSEMBLE && state->ens_spawned)
    rc = quac_mcwf_expect(state, &c, &re, &im);
  else
    rc = qh_expect_wf(state->data, state->total_levels, c.nops, c.ops, c.nlev, c.lev, &re, &im);
  qc_chain_free(&c);
  if (rc) QUAC_FAIL(9, "expectation value failed: %s", qh_last_error());
  *trace_val = re + im * PETSC_i;
}

/* qvec_utilities.c:1133-1151 */
void get_expectation_value_qvec(qvec state, PetscScalar *trace_val, PetscInt num_ops, ...) {
  va_list ap;
  va_start(ap, num_ops);
  operator *ops = (operator *)malloc(sizeof(operator) * (num_ops > 0 ? num_ops : 1));
  for (PetscInt i = 0; i < num_ops; i++) ops[i] = va_arg(ap, operator);
  va_end(ap);
  get_expectation_value_qvec_list(state, trace_val, num_ops, ops);
  free(ops);
}

/* qvec_utilities.c:378-392 */
void get_trace_qvec(qvec dm, PetscScalar *val) {
  if (dm->my_type != DENSITY_MATRIX) QUAC_FAIL(0, "Can only take the trace of density matrices!");
  double re, im;
  QUAC_HIPCHK(qh_trace(dm->data, dm->total_levels, &re, &im));
  *val = re + im * PETSC_i;
}

/* qvec_utilities.c:322-337: loc = sqrt(n)*col + row */
void get_dm_element_qvec(qvec dm, PetscInt row, PetscInt col, PetscScalar *val) {
  double re, im;
  QUAC_HIPCHK(qh_vec_get_entry(dm->data, dm->total_levels * col + row, &re, &im));
  *val = re + im * PETSC_i;
}

void get_wf_element_qvec(qvec wf, PetscInt i, PetscScalar *val) {
  double re, im;
  QUAC_HIPCHK(qh_vec_get_entry(wf->data, i, &re, &im));
  *val = re + im * PETSC_i;
}

/* qvec_utilities.c:1682-1717 (DM: Re diag(rho); WF: |psi_i|^2), arrays malloc'd */
/* a spawned MCWF ensemble (qvec_utilities.c:2041-2086): the mean over the
 * trajectories of |psi_e,i|^2 / <psi_e|psi_e>, and its sample variance */
static void bitstring_probs_ens(qvec state, PetscInt *nloc, PetscReal **probs, PetscReal **vars) {
  const PetscInt N = state->n, M = state->n_ensemble;
  double *h = (double *)malloc(sizeof(double) * 2 * (size_t)(N * M));
  double *nrm = (double *)malloc(sizeof(double) * (size_t)M);
  if (!h || !nrm || qh_vec_get(state->ens_datas[0], h)) {
    free(h);
    free(nrm);
    QUAC_FAIL(9, "get_bitstring_probs: %s", qh_last_error());
  }
  *nloc = N;
  *probs = (PetscReal *)calloc((size_t)N, sizeof(PetscReal));
  *vars = (PetscReal *)calloc((size_t)N, sizeof(PetscReal));
  for (PetscInt e = 0; e < M; e++) {
    const double *p = h + 2 * (size_t)(e * N);
    double s = 0.0;
    for (PetscInt i = 0; i < N; i++) s += p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1];
    nrm[e] = s;
    for (PetscInt i = 0; i < N; i++) (*probs)[i] += (p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1]) / s;
  }
  for (PetscInt i = 0; i < N; i++) (*probs)[i] /= (double)M;
  for (PetscInt e = 0; e < M; e++) {
    const double *p = h + 2 * (size_t)(e * N);
    for (PetscInt i = 0; i < N; i++) {
      const double d = (p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1]) / nrm[e] - (*probs)[i];
      (*vars)[i] += d * d;
    }
  }
  for (PetscInt i = 0; i < N; i++) (*vars)[i] /= (double)(M > 1 ? M - 1 : 1);
  free(h);
  free(nrm);
}

void get_bitstring_probs(qvec state, PetscInt *nloc, PetscReal **probs, PetscReal **vars) {
  if (state->my_type == WF_ENSEMBLE && state->ens_spawned) {
    bitstring_probs_ens(state, nloc, probs, vars);
    return;
  }
  PetscInt N = state->total_levels;
  double *h = (double *)malloc(sizeof(double) * 2 * state->n);
  if (qh_vec_get(state->data, h)) {
    free(h);
    QUAC_FAIL(9, "get_bitstring_probs: %s", qh_last_error());
  }
  *nloc = N;
  *probs = (PetscReal *)malloc(sizeof(PetscReal) * N);
  *vars = (PetscReal *)calloc(N, sizeof(PetscReal));
  for (PetscInt i = 0; i < N; i++) {
    if (state->my_type == DENSITY_MATRIX) (*probs)[i] = h[2 * (N * i + i)];
    else (*probs)[i] = h[2 * i] * h[2 * i] + h[2 * i + 1] * h[2 * i + 1];
  }
  free(h);
}

/* PETSc binary Vec reader (qvec_utilities.c:97-166): {1211214, N} + N complex128, big endian */
static uint64_t be64(const unsigned char *p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v = (v << 8) | p[i];
  return v;
}

static void read_binary(qvec *out, const char *fname, qvec_type type) {
  FILE *f = fopen(fname, "rb");
  if (!f) QUAC_FAIL(9, "cannot open %s", fname);
  unsigned char hdr[16];
  if (fread(hdr, 1, 16, f) != 16 || be64(hdr) != 1211214) {
    fclose(f);
    QUAC_FAIL(9, "%s is not a PETSc binary Vec", fname);
  }
  PetscInt n = (PetscInt)be64(hdr + 8);
  unsigned char *buf = (unsigned char *)malloc((size_t)n * 16);
  size_t got = fread(buf, 1, (size_t)n * 16, f);
  fclose(f);
  if (got != (size_t)n * 16) {
    free(buf);
    QUAC_FAIL(9, "%s is truncated", fname);
  }
  double *h = (double *)malloc(sizeof(double) * 2 * n);
  for (PetscInt i = 0; i < 2 * n; i++) {
    uint64_t v = be64(buf + 8 * i);
    memcpy(&h[i], &v, 8);
  }
  free(buf);
  quac_clear_error();
  create_arb_qvec(out, n, type);
  if (!quac_last_error()[0] && *out) {
    if (qh_vec_set((*out)->data, h)) quac_fail(9, "%s", qh_last_error());
  }
  free(h);
}

void read_qvec_dm_binary(qvec *out, const char fname[]) { read_binary(out, fname, DENSITY_MATRIX); }
void read_qvec_wf_binary(qvec *out, const char fname[]) { read_binary(out, fname, WAVEFUNCTION); }

/* qvec_utilities.c:31-50: same type (and size), or a wavefunction into a
 * density matrix (|psi><psi|) */
void copy_qvec(qvec src, qvec dst) {
  if (src->my_type == WAVEFUNCTION && dst->my_type == DENSITY_MATRIX) {
    copy_qvec_wf_to_dm(src, dst);
    return;
  }
  if (src->my_type == DENSITY_MATRIX && dst->my_type == WAVEFUNCTION)
    QUAC_FAIL(1, "Coping a DM to a WF does not make sense!");
  if (src->n != dst->n) QUAC_FAIL(1, "qvecs have different sizes!");
  QUAC_HIPCHK(qh_vec_copy(src->data, dst->data));
}

void print_qvec(qvec state) {
  double *h = (double *)malloc(sizeof(double) * 2 * state->n);
  if (qh_vec_get(state->data, h)) {
    free(h);
    QUAC_FAIL(9, "print_qvec: %s", qh_last_error());
  }
  if (state->my_type == DENSITY_MATRIX) {
    PetscInt N = state->total_levels;
    for (PetscInt i = 0; i < N; i++) {
      for (PetscInt j = 0; j < N; j++) printf("%4.3e + %4.3ei  ", h[2 * (N * j + i)], h[2 * (N * j + i) + 1]);
      printf("\n");
    }
  } else {
    for (PetscInt i = 0; i < state->n; i++) printf("%4.3e + %4.3ei\n", h[2 * i], h[2 * i + 1]);
  }
  printf("\n");
  free(h);
}

/* accessors for FFI callers */
Mat quac_system_matrix(qsystem s) { return s->mat_allocated ? s->mat_A : NULL; }
Mat quac_system_td_matrix(qsystem s, PetscInt k) {
  return (k >= 0 && k < s->num_time_dep) ? s->time_dep[k].mat_A : NULL;
}
PetscInt quac_system_total_levels(qsystem s) { return s->total_levels; }
PetscInt quac_system_dim(qsystem s) { return s->dim; }
PetscBool quac_system_dm(qsystem s) { return s->dm_equations; }
Vec quac_qvec_data(qvec q) { return q->data; }
void quac_qvec_info(qvec q, PetscInt *n, PetscInt *tl, qvec_type *type) {
  if (n) *n = q->n;
  if (tl) *tl = q->total_levels;
  if (type) *type = q->my_type;
}

/* ------------------------------------------------------- file writers */
/* text dumps (qvec_utilities.c:203-275): "%40.30e  %40.30e" per element, a DM
 * row per line followed by an empty line */
void print_dm_qvec_file(qvec dm, const char filename[]) {
  FILE *fp = fopen(filename, "w");
  if (!fp) QUAC_FAIL(1, "Cannot open file in print_dm_qvec_file= %s", filename);
  double *h = (double *)malloc(sizeof(double) * 2 * dm->n);
  if (qh_vec_get(dm->data, h)) {
    free(h);
    fclose(fp);
    QUAC_FAIL(9, "print_dm_qvec_file: %s", qh_last_error());
  }
  PetscInt N = (PetscInt)llround(sqrt((double)dm->n));
  for (PetscInt i = 0; i < N; i++) {
    for (PetscInt j = 0; j < N; j++) fprintf(fp, "%40.30e  %40.30e ", h[2 * (N * j + i)], h[2 * (N * j + i) + 1]);
    fprintf(fp, "\n");
  }
  fprintf(fp, "\n");
  free(h);
  fclose(fp);
}

void print_wf_qvec_file(qvec state, const char filename[]) {
  FILE *fp = fopen(filename, "w");
  if (!fp) QUAC_FAIL(1, "Cannot open file in print_dm_qvec_file= %s", filename);
  double *h = (double *)malloc(sizeof(double) * 2 * state->n);
  if (qh_vec_get(state->data, h)) {
    free(h);
    fclose(fp);
    QUAC_FAIL(9, "print_wf_qvec_file: %s", qh_last_error());
  }
  for (PetscInt i = 0; i < state->n; i++) fprintf(fp, "%40.30e  %40.30e\n", h[2 * i], h[2 * i + 1]);
  free(h);
  fclose(fp);
}

void print_qvec_file(qvec state, const char filename[]) {
  if (state->my_type == DENSITY_MATRIX)
    print_dm_qvec_file(state, filename);
  else
    print_wf_qvec_file(state, filename);
}

/* PETSc binary formats as written by MatView/VecView on a binary viewer with
 * 64-bit PetscInt and complex scalars (the format of the reference's fixtures,
 * tests/t_helpers.c:7-13): big-endian int64 {1211216, M, N, nz}, M row lengths,
 * nz column indices, nz complex128 values; vectors {1211214, N} + N complex128. */
static void put_be64(FILE *fp, uint64_t v) {
  unsigned char b[8];
  for (int i = 0; i < 8; i++) b[i] = (unsigned char)(v >> (56 - 8 * i));
  fwrite(b, 1, 8, fp);
}
static void put_be_double(FILE *fp, double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  put_be64(fp, u);
}

int quac_write_mat_petsc_binary(Mat A, const char filename[]) {
  int64_t rows = 0, nnz = 0;
  if (!A || qh_mat_info(A, &rows, &nnz)) return quac_fail(9, "quac_write_mat_petsc_binary: no matrix");
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (rows + 1));
  int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
  double *val = (double *)malloc(sizeof(double) * 2 * (nnz > 0 ? nnz : 1));
  int rc = qh_mat_export(A, rp, col, val);
  FILE *fp = rc ? NULL : fopen(filename, "wb");
  if (!rc && !fp) rc = quac_fail(1, "cannot open %s for writing", filename);
  if (fp) {
    put_be64(fp, 1211216);
    put_be64(fp, (uint64_t)rows);
    put_be64(fp, (uint64_t)rows);
    put_be64(fp, (uint64_t)nnz);
    for (int64_t r = 0; r < rows; r++) put_be64(fp, (uint64_t)(rp[r + 1] - rp[r]));
    for (int64_t k = 0; k < nnz; k++) put_be64(fp, (uint64_t)col[k]);
    for (int64_t k = 0; k < 2 * nnz; k++) put_be_double(fp, val[k]);
    if (fclose(fp)) rc = quac_fail(1, "write error on %s", filename);
  }
  free(rp);
  free(col);
  free(val);
  return rc;
}

int quac_write_qvec_petsc_binary(qvec state, const char filename[]) {
  double *h = (double *)malloc(sizeof(double) * 2 * state->n);
  int rc = qh_vec_get(state->data, h);
  FILE *fp = rc ? NULL : fopen(filename, "wb");
  if (rc) rc = quac_fail(9, "quac_write_qvec_petsc_binary: %s", qh_last_error());
  else if (!fp) rc = quac_fail(1, "cannot open %s for writing", filename);
  if (fp) {
    put_be64(fp, 1211214);
    put_be64(fp, (uint64_t)state->n);
    for (PetscInt i = 0; i < 2 * state->n; i++) put_be_double(fp, h[i]);
    if (fclose(fp)) rc = quac_fail(1, "write error on %s", filename);
  }
  free(h);
  return rc;
}
