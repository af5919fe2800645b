/* quac_core.c -- initialisation, options, error conventions, op descriptors.
 * QuaC_initialize/QuaC_finalize follow src/quac.c:25-141 (device set-up in
 * place of SLEPc/PETSc/MPI initialisation). */
#include <math.h>
#include <string.h>

#include "quac_internal.h"

static int g_initialized = 0;
static int g_err_mode = QUAC_ERRORS_EXIT;
static char g_err[1024] = "";

int PetscPrintf(int comm, const char *fmt, ...) {
  (void)comm;
  va_list ap;
  va_start(ap, fmt);
  int r = vprintf(fmt, ap);
  va_end(ap);
  return r;
}

void quac_set_error_mode(int mode) { g_err_mode = mode; }
int quac_errors_return(void) { return g_err_mode == QUAC_ERRORS_RETURN; }
const char *quac_last_error(void) { return g_err; }
void quac_clear_error(void) { g_err[0] = 0; }

int quac_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  if (g_err_mode == QUAC_ERRORS_EXIT) {
    fprintf(stdout, "ERROR! %s\n", g_err);
    fflush(stdout);
    exit(code);
  }
  return code;
}

/* ---------------------------------------------------------------- options */
#define MAXOPT 64
static char g_opt_name[MAXOPT][64];
static char g_opt_val[MAXOPT][64];
static int g_nopt = 0;

void quac_set_option(const char *name, const char *value) {
  if (!name) return;
  if (name[0] == '-') name++;
  for (int i = 0; i < g_nopt; i++)
    if (!strcmp(g_opt_name[i], name)) {
      snprintf(g_opt_val[i], sizeof(g_opt_val[i]), "%s", value ? value : "");
      return;
    }
  if (g_nopt == MAXOPT) return;
  snprintf(g_opt_name[g_nopt], sizeof(g_opt_name[0]), "%s", name);
  snprintf(g_opt_val[g_nopt], sizeof(g_opt_val[0]), "%s", value ? value : "");
  g_nopt++;
}

void quac_unset_option(const char *name) {
  if (name && name[0] == '-') name++;
  for (int i = 0; i < g_nopt; i++)
    if (!strcmp(g_opt_name[i], name)) {
      g_nopt--;
      if (i != g_nopt) {
        memcpy(g_opt_name[i], g_opt_name[g_nopt], sizeof(g_opt_name[0]));
        memcpy(g_opt_val[i], g_opt_val[g_nopt], sizeof(g_opt_val[0]));
      }
      return;
    }
}

const char *quac_get_option(const char *name) {
  if (name && name[0] == '-') name++;
  for (int i = 0; i < g_nopt; i++)
    if (!strcmp(g_opt_name[i], name)) return g_opt_val[i];
  return NULL;
}

const char *quac_opt_str(const char *name, const char *dflt) {
  const char *v = quac_get_option(name);
  return v ? v : dflt;
}
int quac_opt_int(const char *name, int dflt) {
  const char *v = quac_get_option(name);
  return v && *v ? atoi(v) : dflt;
}
double quac_opt_double(const char *name, double dflt) {
  const char *v = quac_get_option(name);
  return v && *v ? atof(v) : dflt;
}

/* -name value pairs from argv (the PETSc options-database convention) */
static void parse_args(int argc, char **args) {
  for (int i = 1; i < argc; i++) {
    if (!args[i] || args[i][0] != '-') continue;
    const char *val = (i + 1 < argc && args[i + 1] && args[i + 1][0] != '-') ? args[i + 1] : "1";
    quac_set_option(args[i], val);
  }
}

int quac_ensure_device(void) {
  if (!g_initialized) {
    int rc = qh_init(0);
    if (rc) return quac_fail(9, "no usable HIP device: %s", qh_last_error());
    g_initialized = 1;
  }
  return 0;
}

void QuaC_initialize(int argc, char **args) {
  parse_args(argc, args);
  const char *env = getenv("QUAC_OPTIONS");
  if (env) {
    char buf[1024];
    snprintf(buf, sizeof(buf), "%s", env);
    char *toks[64];
    int n = 1;
    toks[0] = "quac";
    for (char *t = strtok(buf, " "); t && n < 64; t = strtok(NULL, " ")) toks[n++] = t;
    parse_args(n, toks);
  }
  if (quac_ensure_device()) return;
}

void QuaC_clear(void) { qh_synchronize(); }

void QuaC_finalize(void) {
  time_step_sys_batch_release();
  if (g_initialized) qh_finalize();
  g_initialized = 0;
}

/* ------------------------------------------------------ chain descriptors */
void qc_chain_init(qc_chain *c) { memset(c, 0, sizeof(*c)); }
void qc_chain_free(qc_chain *c) {
  free(c->ops);
  free(c->lev);
  memset(c, 0, sizeof(*c));
}

static qh_level *lev_alloc(qc_chain *c, int n) {
  if (c->nlev + n > c->cap_lev) {
    int cap = c->cap_lev ? 2 * c->cap_lev : 64;
    while (cap < c->nlev + n) cap *= 2;
    c->lev = (qh_level *)realloc(c->lev, sizeof(qh_level) * (size_t)cap);
    c->cap_lev = cap;
  }
  qh_level *p = c->lev + c->nlev;
  for (int i = 0; i < n; i++) {
    p[i].next = -1;
    p[i].prev = -1;
    p[i].re = 0.0;
    p[i].im = 0.0;
  }
  c->nlev += n;
  return p;
}

static void op_push(qc_chain *c, int64_t stride, int levels, int tab) {
  if (c->nops == c->cap_ops) {
    c->cap_ops = c->cap_ops ? 2 * c->cap_ops : 16;
    c->ops = (qh_op *)realloc(c->ops, sizeof(qh_op) * (size_t)c->cap_ops);
  }
  c->ops[c->nops].stride = stride;
  c->ops[c->nops].levels = levels;
  c->ops[c->nops].tab = tab;
  c->nops++;
}

static void set_map(qh_level *t, int from, int to, double re, double im) {
  t[from].next = to;
  t[from].re = re;
  t[from].im = im;
  t[to].prev = from;
}

/* per-level rule of kron.c:294-538 written as a table */
int qc_chain_append(qc_chain *c, PetscInt tot, PetscInt nops, operator *ops) {
  int appended = 0;
  for (PetscInt k = 0; k < nops; k++) {
    operator op = ops[k];
    if (!op) return quac_fail(0, "null operator in term"), -1;
    const int L = (int)op->my_levels;
    if (L <= 0 || tot % (op->my_levels * op->n_before) != 0)
      return quac_fail(0, "operator does not belong to this Hilbert space"), -1;
    const int64_t stride = tot / (op->my_levels * op->n_before);
    const int tab = c->nlev;
    qh_level *t = lev_alloc(c, L);
    switch (op->my_op_type) {
      case LOWER:
        for (int l = 0; l < L - 1; l++) set_map(t, l, l + 1, sqrt((double)l + 1.0), 0.0);
        break;
      case RAISE:
        for (int l = 1; l < L; l++) set_map(t, l, l - 1, sqrt((double)(l - 1) + 1.0), 0.0);
        break;
      case NUMBER:
        for (int l = 1; l < L; l++) set_map(t, l, l, (double)l, 0.0);
        break;
      case IDENTITY:
        for (int l = 0; l < L; l++) set_map(t, l, l, 1.0, 0.0);
        break;
      case SIGMA_X:
        if (L >= 2) {
          set_map(t, 0, 1, 1.0, 0.0);
          set_map(t, 1, 0, 1.0, 0.0);
        }
        break;
      case SIGMA_Y:
        if (L != 2) return quac_fail(0, "Pauli Operators are only defined for qubits"), -1;
        set_map(t, 0, 1, -0.0, -1.0); /* -PETSC_i */
        set_map(t, 1, 0, 0.0, 1.0);
        break;
      case SIGMA_Z:
        if (L != 2) return quac_fail(0, "Pauli Operators are only defined for qubits"), -1;
        set_map(t, 0, 0, 1.0, 0.0);
        set_map(t, 1, 1, -1.0, 0.0);
        break;
      case OTHER:
        for (int l = 2; l < L; l++) set_map(t, l, l, 1.0, 0.0);
        break;
      case VEC: {
        /* VEC operators come in pairs |p1><p2| (kron.c:170-190) */
        if (k + 1 >= nops || ops[k + 1]->my_op_type != VEC)
          return quac_fail(0, "VEC operators must come in pairs"), -1;
        operator op2 = ops[k + 1];
        if (op2->n_before != op->n_before)
          return quac_fail(0, "Vec Operators must be from the same Hilbert subspace"), -1;
        set_map(t, (int)op->position, (int)op2->position, 1.0, 0.0);
        k++;
        break;
      }
      default:
        return quac_fail(0, "unknown operator type %d", (int)op->my_op_type), -1;
    }
    op_push(c, stride, L, tab);
    appended++;
  }
  return appended;
}

operator qc_dag(operator op) { return op->my_op_type == VEC ? op : op->dag; }
