/* quac_internal.h -- shared helpers of the C host layer (not installed). */
#ifndef QUAC_INTERNAL_H
#define QUAC_INTERNAL_H

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "quac.h"

/* error handling: print "ERROR! ..." then exit(code) (reference behaviour) or
 * record the message and let the caller return (QUAC_ERRORS_RETURN). */
int quac_fail(int code, const char *fmt, ...);
int quac_errors_return(void);
#define QUAC_FAIL(code, ...)          \
  do {                                \
    quac_fail((code), __VA_ARGS__);   \
    return;                           \
  } while (0)
#define QUAC_FAIL_RC(rc, code, ...)   \
  do {                                \
    quac_fail((code), __VA_ARGS__);   \
    return (rc);                      \
  } while (0)
/* map a qh_* failure to the error path */
#define QUAC_HIPCHK(call)                                                     \
  do {                                                                        \
    int rc_ = (call);                                                         \
    if (rc_ != QH_OK) QUAC_FAIL(9, "device call %s failed: %s", #call, qh_last_error()); \
  } while (0)
#define QUAC_HIPCHK_RC(call, ret)                                             \
  do {                                                                        \
    int rc_ = (call);                                                         \
    if (rc_ != QH_OK) QUAC_FAIL_RC(ret, 9, "device call %s failed: %s", #call, qh_last_error()); \
  } while (0)

int quac_ensure_device(void);

/* MCWF trajectories (mcwf.c) */
void quac_rng_seed(qsystem s, uint64_t seed);
/* the same generator on a caller's 4-word state (seed 0: from the clock) */
void quac_rng_seed_state(unsigned long long *q, uint64_t seed);
double quac_rng_uniform_state(unsigned long long *q);
double quac_rng_uniform(qsystem s);
void quac_mcwf_free(qsystem s);
/* the ensemble from t0 to tmax (steps of dt, the last one cut at tmax, at most
 * steps_max); the monitor sees step numbers from step0, and the state at t0
 * only when call_initial */
int quac_mcwf_time_step(qsystem s, qvec x, PetscReal t0, PetscReal tmax, PetscReal dt, PetscInt steps_max,
                        int (*monitor)(qsystem, PetscInt, PetscReal, qvec), PetscInt *nsteps, PetscInt step0,
                        int call_initial);

/* Wootters' concurrence of a two-qubit density matrix (4 x 4, column-major) */
int quac_concurrence_2q(const double _Complex *rho, double *conc);

/* options (quac_unset_option: forget a name set by quac_set_option) */
void quac_unset_option(const char *name);
int quac_opt_int(const char *name, int dflt);
double quac_opt_double(const char *name, double dflt);
const char *quac_opt_str(const char *name, const char *dflt);

/* growable device-assembly descriptor lists */
typedef struct {
  qh_op *ops;
  int nops, cap_ops;
  qh_level *lev;
  int nlev, cap_lev;
} qc_chain;
void qc_chain_init(qc_chain *c);
void qc_chain_free(qc_chain *c);
int quac_mcwf_expect(qvec x, qc_chain *c, double *re, double *im);
/* Append the ops of a QuaC operator product to c (a consecutive pair of VEC
 * operators becomes one |p1><p2| entry).  Returns number of appended entries,
 * or -1 on error (message recorded).  tot = total_levels of the space. */
int qc_chain_append(qc_chain *c, PetscInt tot, PetscInt nops, operator *ops);

/* dagger of a basic operator (LOWER<->RAISE, others self) */
operator qc_dag(operator op);

/* circuit events during time stepping (circuits.c) */
int qc_next_event_time(qsystem sys, PetscReal *t_event);
/* applies the gates (or layer) due at the current event; nonzero on failure */
int qc_apply_events(qsystem sys, PetscReal t, qvec x);
/* device gate matrix and strides for a gate; superop=1: qubit q -> subsystem q
 * (stride n_after), else qubit q -> stride prod(dims[0..q-1]) */
int qc_apply_gate(struct quantum_gate_struct *g, qvec x, int superop, qsystem sys);

#endif
