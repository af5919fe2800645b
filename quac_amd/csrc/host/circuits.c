/* circuits.c -- circuits, gate scheduling, gate events and gate application of
 * the QuaC-compatible host layer (src/quantum_circuits.c, src/quantum_gates.c).
 *
 * Two qubit conventions exist in the reference and both are kept:
 *  - superoperator path (_apply_gate_sys, quantum_circuits.c:435-481; used by
 *    apply_circuit_to_qvec and scheduled layers): qubit q is subsystem q;
 *  - in-place path (_apply_gate2, :696-772; used by apply_gate2 and unscheduled
 *    circuit events): qubit q has stride prod(dims[0..q-1]).
 * Both end in one device pass rho <- U rho U^dagger (qh_gate).  RZ is applied
 * as U rho U^dagger on both paths; RZ_gate_func_wf (:2114-2147) ignores the
 * conjugation on the column side, which this library does not reproduce.
 */
#include <math.h>
#include <string.h>

#include "quac_internal.h"

/* _check_gate_type, quantum_gates.c:3001-3017 */
static int gate_nqubits(gate_type t) {
  switch (t) {
    case HADAMARD: case SIGMAX: case SIGMAY: case SIGMAZ: case EYE: case RX: case RY: case RZ: case U1: case U2:
    case U3: case CUSTOM1QGATE:
      return 1;
    case CNOT: case CXZ: case CZ: case CmZ: case CZX: case CUSTOM2QGATE: case CZ_ARP: case SQiSWP:
      return 2;
    default:
      return 0;
  }
}

/* quantum_gates.c:610-628 */
void create_circuit(circuit *circ, PetscInt num_gates) {
  if (num_gates < 1) num_gates = 1;
  circ->num_gates = 0;
  circ->gate_list_size = num_gates;
  circ->current_gate = 0;
  circ->num_layers = 0;
  circ->current_layer = 0;
  circ->start_time = 0.0;
  circ->gate_list = (struct quantum_gate_struct *)calloc(num_gates, sizeof(struct quantum_gate_struct));
  circ->layer_list = (struct gate_layer_struct *)calloc(num_gates, sizeof(struct gate_layer_struct));
}

/* quantum_gates.c:630-649 */
void destroy_circuit(circuit *circ) {
  for (PetscInt i = 0; i < circ->num_gates; i++) free(circ->gate_list[i].qubit_numbers);
  for (PetscInt i = 0; i < circ->gate_list_size; i++)
    if (circ->layer_list) free(circ->layer_list[i].gate_list);
  free(circ->gate_list);
  free(circ->layer_list);
  circ->gate_list = NULL;
  circ->layer_list = NULL;
  circ->num_gates = 0;
}

static void fill_gate(struct quantum_gate_struct *g, PetscReal time, gate_type type, va_list ap) {
  int nq = gate_nqubits(type);
  g->qubit_numbers = (PetscInt *)malloc(sizeof(PetscInt) * (nq > 0 ? nq : 1));
  g->time = time;
  g->run_time = -1;
  g->my_gate_type = type;
  g->num_qubits = nq;
  for (int i = 0; i < nq; i++) g->qubit_numbers[i] = va_arg(ap, int);
  g->theta = g->phi = g->lambda = 0.0;
  g->gate_ctx = NULL;
  g->custom_func = NULL;
  if (type == RX || type == RY || type == RZ) {
    g->theta = va_arg(ap, PetscReal);
  } else if (type == U1 || type == U2 || type == U3) {
    g->theta = va_arg(ap, PetscReal);
    g->phi = va_arg(ap, PetscReal);
    g->lambda = va_arg(ap, PetscReal);
  } else if (type == CUSTOM1QGATE || type == CUSTOM2QGATE) {
    g->gate_ctx = va_arg(ap, void *);
    g->custom_func = va_arg(ap, custom_gate_func_type);
  }
}

/* quantum_circuits.c:172-255 */
void add_gate_to_circuit_sys(circuit *circ, PetscReal time, gate_type type, ...) {
  if (gate_nqubits(type) == 0) QUAC_FAIL(0, "Gate not recognized!");
  if (circ->num_gates == circ->gate_list_size) QUAC_FAIL(1, "Gate list not large enough!");
  va_list ap;
  va_start(ap, type);
  fill_gate(&circ->gate_list[circ->num_gates], time, type, ap);
  va_end(ap);
  circ->num_gates++;
}

/* quantum_gates.c:651-714: the same gate record for the legacy global API (its
 * circuits are started with start_circuit_at_time) */
void add_gate_to_circuit(circuit *circ, PetscReal time, gate_type type, ...) {
  if (gate_nqubits(type) == 0) QUAC_FAIL(0, "Gate not recognized!");
  if (circ->num_gates == circ->gate_list_size) QUAC_FAIL(1, "Gate list not large enough!");
  va_list ap;
  va_start(ap, type);
  fill_gate(&circ->gate_list[circ->num_gates], time, type, ap);
  va_end(ap);
  circ->num_gates++;
}

/* quantum_gates.c:716-749: append every gate of circ_to_add to circ at `time` */
void add_circuit_to_circuit(circuit *circ, circuit circ_to_add, PetscReal time) {
  if (circ->num_gates + circ_to_add.num_gates > circ->gate_list_size)
    QUAC_FAIL(1, "Gate list not large enough to add this circuit!");
  for (PetscInt i = 0; i < circ_to_add.num_gates; i++) {
    struct quantum_gate_struct *g = &circ->gate_list[circ->num_gates];
    *g = circ_to_add.gate_list[i];
    g->time = time;
    g->qubit_numbers = (PetscInt *)malloc(sizeof(PetscInt) * (size_t)(g->num_qubits > 0 ? g->num_qubits : 1));
    for (PetscInt j = 0; j < g->num_qubits; j++) g->qubit_numbers[j] = circ_to_add.gate_list[i].qubit_numbers[j];
    circ->num_gates++;
  }
}

/* quantum_circuits.c:316-404: greedy ASAP layering */
void schedule_circuit_layers(qsystem sys, circuit *circ) {
  PetscInt ns = sys->num_subsystems;
  PetscInt *cur = (PetscInt *)calloc(ns > 0 ? ns : 1, sizeof(PetscInt));
  for (PetscInt i = 0; i < circ->num_gates; i++) {
    free(circ->layer_list[i].gate_list);
    circ->layer_list[i].gate_list =
        (struct quantum_gate_struct *)calloc(ns > 0 ? ns : 1, sizeof(struct quantum_gate_struct));
    circ->layer_list[i].num_gates = 0;
  }
  for (PetscInt i = 0; i < circ->num_gates; i++) {
    struct quantum_gate_struct *g = &circ->gate_list[i];
    PetscInt maxl = 0;
    for (PetscInt j = 0; j < g->num_qubits; j++) {
      PetscInt q = g->qubit_numbers[j];
      if (q < 0 || q >= ns) {
        free(cur);
        QUAC_FAIL(0, "gate qubit %lld is not a subsystem of this system", (long long)q);
      }
      if (cur[q] > maxl) maxl = cur[q];
    }
    struct gate_layer_struct *L = &circ->layer_list[maxl];
    L->gate_list[L->num_gates++] = *g;
    for (PetscInt j = 0; j < g->num_qubits; j++) cur[g->qubit_numbers[j]] = maxl + 1;
  }
  PetscInt maxl = 0;
  for (PetscInt j = 0; j < ns; j++)
    if (cur[j] > maxl) maxl = cur[j];
  circ->num_layers = maxl;
  if (maxl > 0 && circ->layer_list[0].num_gates > 0 && circ->layer_list[0].gate_list[0].run_time > 0) {
    PetscReal t = 1.0;
    for (PetscInt i = 0; i < circ->num_layers; i++) {
      circ->layer_list[i].time = t;
      PetscReal mx = -2;
      for (PetscInt j = 0; j < circ->layer_list[i].num_gates; j++)
        if (circ->layer_list[i].gate_list[j].run_time > mx) mx = circ->layer_list[i].gate_list[j].run_time;
      if (mx < 0) {
        PetscPrintf(0, "WARNING: All gate run times for layer were negative; defaulting to 1.0\n");
        mx = 1.0;
      }
      t += mx;
    }
  } else {
    for (PetscInt i = 0; i < circ->num_layers; i++) circ->layer_list[i].time = (PetscReal)(i + 1);
  }
  free(cur);
}

/* quantum_circuits.c:407-431 (stores a shallow copy, sharing gate_list) */
void apply_circuit_to_sys(qsystem sys, circuit *circ, PetscReal time) {
  circ->start_time = time;
  if (sys->num_circuits == -1) {
    sys->circuit_list_size = 10;
    sys->circuit_list = (circuit *)malloc(sys->circuit_list_size * sizeof(circuit));
    sys->num_circuits = 0;
    sys->current_circuit = 0;
  }
  if (sys->num_circuits == sys->circuit_list_size) QUAC_FAIL(1, "Circuit list not large enough!");
  sys->circuit_list[sys->num_circuits++] = *circ;
}

/* ----------------------------------------------------------- gate math */
static void u2_of(const struct quantum_gate_struct *g, double u[8]) {
  const double s2 = pow(2.0, -0.5);
  double th = g->theta, ph = g->phi, la = g->lambda;
  double m[8] = {0};
  switch (g->my_gate_type) {
    case HADAMARD: m[0] = s2; m[2] = s2; m[4] = s2; m[6] = -s2; break;
    case SIGMAX: m[2] = 1; m[4] = 1; break;
    case SIGMAY: m[3] = -1; m[5] = 1; break;
    case SIGMAZ: m[0] = 1; m[6] = -1; break;
    case EYE: m[0] = 1; m[6] = 1; break;
    case RX:  /* quantum_circuits.c:2650-2739 */
      m[0] = cos(th / 2); m[3] = -sin(th / 2); m[5] = -sin(th / 2); m[6] = cos(th / 2); break;
    case RY: m[0] = cos(th / 2); m[2] = -sin(th / 2); m[4] = sin(th / 2); m[6] = cos(th / 2); break;
    case RZ:  /* diag(1, e^{i theta}), :2484-2558 */
      m[0] = 1; m[6] = cos(th); m[7] = sin(th); break;
    case U1: case U2: case U3: {  /* U3_gate_func_wf, :2055-2110 */
      double c = cos(th / 2), s = sin(th / 2);
      m[0] = c;
      m[2] = -cos(la) * s; m[3] = -sin(la) * s;
      m[4] = cos(ph) * s; m[5] = sin(ph) * s;
      m[6] = cos(la + ph) * c; m[7] = sin(la + ph) * c;
      break;
    }
    case CUSTOM1QGATE: {
      custom_gate_data *d = (custom_gate_data *)g->gate_ctx;
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          m[2 * (2 * i + j)] = creal(d->gate_data[i][j]);
          m[2 * (2 * i + j) + 1] = cimag(d->gate_data[i][j]);
        }
      break;
    }
    default: break;
  }
  memcpy(u, m, sizeof(m));
}

static void u4_of(const struct quantum_gate_struct *g, double u[32]) {
  memset(u, 0, sizeof(double) * 32);
#define SET(i, j, re, im) (u[2 * (4 * (i) + (j))] = (re), u[2 * (4 * (i) + (j)) + 1] = (im))
  const double s2 = 1.0 / sqrt(2.0);
  switch (g->my_gate_type) {
    case CNOT: SET(0, 0, 1, 0); SET(1, 1, 1, 0); SET(2, 3, 1, 0); SET(3, 2, 1, 0); break;
    case CXZ: SET(0, 0, 1, 0); SET(1, 1, 1, 0); SET(2, 3, -1, 0); SET(3, 2, 1, 0); break;
    case CZ: SET(0, 0, 1, 0); SET(1, 1, 1, 0); SET(2, 2, 1, 0); SET(3, 3, -1, 0); break;
    case CmZ: SET(0, 0, 1, 0); SET(1, 1, 1, 0); SET(2, 2, -1, 0); SET(3, 3, 1, 0); break;
    case CZX: SET(0, 0, 1, 0); SET(1, 1, 1, 0); SET(2, 3, 1, 0); SET(3, 2, -1, 0); break;
    case CZ_ARP: SET(0, 0, 1, 0); SET(1, 1, -1, 0); SET(2, 2, -1, 0); SET(3, 3, -1, 0); break;
    case SQiSWP:
      SET(0, 0, 1, 0); SET(1, 1, s2, 0); SET(1, 2, 0, s2); SET(2, 1, 0, s2); SET(2, 2, s2, 0); SET(3, 3, 1, 0);
      break;
    case CUSTOM2QGATE: {
      custom_gate_data *d = (custom_gate_data *)g->gate_ctx;
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) SET(i, j, creal(d->gate_data[i][j]), cimag(d->gate_data[i][j]));
      break;
    }
    default: break;
  }
#undef SET
}

int qc_apply_gate(struct quantum_gate_struct *g, qvec x, int superop, qsystem sys) {
  int nq = gate_nqubits(g->my_gate_type);
  if (nq == 0) return quac_fail(0, "Gate not recognized!");
  if ((g->my_gate_type == CUSTOM1QGATE || g->my_gate_type == CUSTOM2QGATE) && !g->gate_ctx)
    return quac_fail(0, "custom gate without gate data");
  const int dm = x->my_type == DENSITY_MATRIX;
  const PetscInt tot = x->total_levels;
  int64_t strides[2];
  for (int q = 0; q < nq; q++) {
    PetscInt qb = g->qubit_numbers[q];
    if (superop) {
      if (!sys || qb < 0 || qb >= sys->num_subsystems) return quac_fail(0, "qubit %lld out of range", (long long)qb);
      operator op = sys->subsystem_list[qb];
      if (op->my_levels != 2) return quac_fail(0, "Single qubit gates can only affect 2-level systems");
      strides[q] = sys->total_levels / (op->my_levels * op->n_before);
    } else {
      PetscInt ns = dm ? x->ndims_hspace / 2 : x->ndims_hspace;
      if (qb < 0 || qb >= ns) return quac_fail(0, "qubit %lld out of range", (long long)qb);
      if (x->hspace_dims[qb] != 2) return quac_fail(0, "Single qubit gates can only affect 2-level systems");
      PetscInt sb = 1;
      for (PetscInt k = 0; k < qb; k++) sb *= x->hspace_dims[k];
      strides[q] = sb;
    }
  }
  double u[32];
  if (nq == 1) u2_of(g, u);
  else {
    u4_of(g, u);
    /* in-place CUSTOM2Q indexes its 4x4 data with the first qubit as the low bit */
    if (!superop && g->my_gate_type == CUSTOM2QGATE) {
      int64_t t = strides[0];
      strides[0] = strides[1];
      strides[1] = t;
    }
  }
  int rc = qh_gate(x->data, tot, dm, nq, strides, u);
  if (rc) return quac_fail(9, "gate application failed: %s", qh_last_error());
  return 0;
}

/* the gate on every trajectory of a spawned MCWF ensemble (quantum_circuits.c:
 * 127-132, 146-151), otherwise on the state itself */
static int apply_on_members(struct quantum_gate_struct *g, qvec state, int superop, qsystem sys) {
  if (state->my_type != WF_ENSEMBLE || !state->ens_spawned) return qc_apply_gate(g, state, superop, sys);
  /* the trajectories are consecutive blocks of N = total_levels entries, and a
   * qubit's stride s has 2s dividing N, so one launch over the whole ensemble as
   * a single vector of M N entries pairs the same entries as M launches would */
  struct qvec all = *state;
  all.my_type = WAVEFUNCTION;
  all.total_levels = state->n * state->n_ensemble;
  Vec v = NULL;
  if (qh_vec_view(state->ens_datas[0], 0, all.total_levels, &v)) return quac_fail(9, "%s", qh_last_error());
  all.data = v;
  const int rc = qc_apply_gate(g, &all, superop, sys);
  qh_vec_destroy(v);
  return rc;
}

/* quantum_circuits.c:257-314: the circuit's matrix G_n ... G_1 on the Hilbert
 * space (total_levels rows; gate times are ignored, as there).  Column j is the
 * circuit applied to the basis state e_j, gate by gate, with the device gate
 * path and qubit convention of apply_circuit_to_qvec on a wavefunction (the
 * reference builds each factor with the same row generators, argument -1).  The
 * matrix is dense on the host while it is built: total_levels <= 2048. */
void combine_circuit_to_mat_sys(qsystem sys, Mat *matrix_out, circuit circ) {
  *matrix_out = NULL;
  if (!sys || sys->num_subsystems == 0) QUAC_FAIL(0, "combine_circuit_to_mat_sys: the system has no subsystems");
  if (circ.num_gates < 1) QUAC_FAIL(0, "combine_circuit_to_mat_sys: the circuit has no gates");
  const PetscInt N = sys->total_levels;
  if (N > 2048) QUAC_FAIL(9, "combine_circuit_to_mat_sys: total_levels %lld is above 2048", (long long)N);
  PetscInt *dims = (PetscInt *)malloc(sizeof(PetscInt) * (size_t)sys->num_subsystems);
  double complex *U = (double complex *)malloc(sizeof(double complex) * (size_t)(N * N));
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
  int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N * N));
  double complex *val = (double complex *)malloc(sizeof(double complex) * (size_t)(N * N));
  qvec q = NULL;
  int rc = !dims || !U || !rp || !col || !val;
  if (!rc) {
    for (PetscInt i = 0; i < sys->num_subsystems; i++) dims[i] = sys->subsystem_list[i]->my_levels;
    create_arb_qvec_dims(&q, sys->num_subsystems, dims, WAVEFUNCTION);
    rc = !q;
  }
  for (PetscInt j = 0; j < N && !rc; j++) {
    rc = qh_vec_zero(q->data) || qh_vec_add_entry(q->data, j, 1.0, 0.0);
    for (PetscInt g = 0; g < circ.num_gates && !rc; g++) rc = qc_apply_gate(&circ.gate_list[g], q, 1, sys);
    if (!rc) rc = qh_vec_get(q->data, (double *)(U + (size_t)N * j));
  }
  if (!rc) {
    int64_t nnz = 0;
    for (PetscInt r = 0; r < N; r++) {
      rp[r] = nnz;
      for (PetscInt c = 0; c < N; c++)
        if (U[(size_t)N * c + r] != 0.0) {
          col[nnz] = c;
          val[nnz++] = U[(size_t)N * c + r];
        }
    }
    rp[N] = nnz;
    rc = qh_mat_from_csr(N, rp, col, (const double *)val, matrix_out);
  }
  if (q) destroy_qvec(&q);
  free(dims);
  free(U);
  free(rp);
  free(col);
  free(val);
  if (rc) QUAC_FAIL(9, "combine_circuit_to_mat_sys: %s", qh_last_error());
}

/* quantum_circuits.c:143-159 */
void apply_circuit_to_qvec(qsystem qsys, circuit circ, qvec state) {
  for (PetscInt i = 0; i < circ.num_gates; i++)
    if (apply_on_members(&circ.gate_list[i], state, 1, qsys)) return;
}

/* quantum_circuits.c:123-140: every gate with the in-place convention of apply_gate2 */
void apply_circuit_to_qvec2(circuit circ, qvec state) {
  for (PetscInt i = 0; i < circ.num_gates; i++)
    if (apply_on_members(&circ.gate_list[i], state, 0, NULL)) return;
}

/* quantum_circuits.c:776-847 */
void apply_gate2(qvec state, gate_type type, ...) {
  if (gate_nqubits(type) == 0) QUAC_FAIL(0, "Gate not recognized in get_gate_func_wf!");
  struct quantum_gate_struct g;
  va_list ap;
  va_start(ap, type);
  fill_gate(&g, 1.0, type, ap);
  va_end(ap);
  qc_apply_gate(&g, state, 0, NULL);
  free(g.qubit_numbers);
}

/* ------------------------------------------------- events during stepping */
/* _sys_QC_EventFunction, quantum_circuits.c:17-56: the next gate/layer time */
int qc_next_event_time(qsystem sys, PetscReal *t) {
  if (sys->num_circuits <= 0 || sys->current_circuit >= sys->num_circuits) return 0;
  circuit *c = &sys->circuit_list[sys->current_circuit];
  if (c->num_layers > 0) {
    if (c->current_layer >= c->num_layers) return 0;
    *t = c->layer_list[c->current_layer].time + c->start_time;
  } else {
    if (c->current_gate >= c->num_gates) return 0;
    *t = c->gate_list[c->current_gate].time + c->start_time;
  }
  return 1;
}

/* _sys_QC_PostEventFunction, quantum_circuits.c:58-120 */
int qc_apply_events(qsystem sys, PetscReal t, qvec x) {
  (void)t;
  if (sys->current_circuit >= sys->num_circuits) return 0;
  circuit *c = &sys->circuit_list[sys->current_circuit];
  if (c->num_layers > 0) {
    PetscReal lt = c->layer_list[c->current_layer].time;
    while (c->current_layer < c->num_layers && c->layer_list[c->current_layer].time == lt) {
      struct gate_layer_struct *L = &c->layer_list[c->current_layer];
      for (PetscInt i = 0; i < L->num_gates; i++)
        if (apply_on_members(&L->gate_list[i], x, 1, sys)) return 1;
      c->current_layer++;
    }
    if (c->current_layer >= c->num_layers) sys->current_circuit++;
  } else {
    PetscReal gt = c->gate_list[c->current_gate].time;
    while (c->current_gate < c->num_gates && c->gate_list[c->current_gate].time == gt) {
      if (apply_on_members(&c->gate_list[c->current_gate], x, 0, sys)) return 1;
      c->current_gate++;
    }
    if (c->current_gate >= c->num_gates) sys->current_circuit++;
  }
  return 0;
}

/* ------------------------------------------- non-variadic forms for FFI */
static void fill_gate_list(struct quantum_gate_struct *g, PetscReal time, gate_type type, const PetscInt *qubits,
                           PetscReal theta, PetscReal phi, PetscReal lambda, void *ctx) {
  int nq = gate_nqubits(type);
  g->qubit_numbers = (PetscInt *)malloc(sizeof(PetscInt) * (nq > 0 ? nq : 1));
  for (int i = 0; i < nq; i++) g->qubit_numbers[i] = qubits[i];
  g->time = time;
  g->run_time = -1;
  g->my_gate_type = type;
  g->num_qubits = nq;
  g->theta = theta;
  g->phi = phi;
  g->lambda = lambda;
  g->gate_ctx = ctx;
  g->custom_func = NULL;
}

void add_gate_to_circuit_sys_list(circuit *circ, PetscReal time, gate_type type, const PetscInt *qubits,
                                  PetscReal theta, PetscReal phi, PetscReal lambda, void *ctx) {
  if (gate_nqubits(type) == 0) QUAC_FAIL(0, "Gate not recognized!");
  if (circ->num_gates == circ->gate_list_size) QUAC_FAIL(1, "Gate list not large enough!");
  fill_gate_list(&circ->gate_list[circ->num_gates], time, type, qubits, theta, phi, lambda, ctx);
  circ->num_gates++;
}

void apply_gate2_list(qvec state, gate_type type, const PetscInt *qubits, PetscReal theta, PetscReal phi,
                      PetscReal lambda, void *ctx) {
  if (gate_nqubits(type) == 0) QUAC_FAIL(0, "Gate not recognized in get_gate_func_wf!");
  struct quantum_gate_struct g;
  fill_gate_list(&g, 1.0, type, qubits, theta, phi, lambda, ctx);
  qc_apply_gate(&g, state, 0, NULL);
  free(g.qubit_numbers);
}

/* one gate through the superoperator convention (_apply_gate_sys) */
void apply_gate_sys_list(qsystem sys, qvec state, gate_type type, const PetscInt *qubits, PetscReal theta,
                         PetscReal phi, PetscReal lambda, void *ctx) {
  if (gate_nqubits(type) == 0) QUAC_FAIL(0, "Gate not recognized!");
  struct quantum_gate_struct g;
  fill_gate_list(&g, 1.0, type, qubits, theta, phi, lambda, ctx);
  qc_apply_gate(&g, state, 1, sys);
  free(g.qubit_numbers);
}
