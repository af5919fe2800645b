/* mcwf.c -- Monte Carlo wavefunction trajectories of a qsystem under
 * use_mcwf_solver (qsystem.c:358-403, 1024-1079, 1138-1235).
 *
 * All trajectories advance together on the device: RK4 steps of
 * d psi/dt = -i Heff psi, and after every step the squared norm of each
 * trajectory is compared with its threshold r_i.  When it fell below r_i during
 * the step, the crossing time inside the step is located (the event location of
 * qsystem.c:1017-1019): the RK4 step is a polynomial in its size on this linear
 * system, so the squared norm is a degree-8 polynomial whose root is found to
 * machine precision, and the state there is the step re-taken to that time.  The
 * trajectory jumps there (operator k with probability a_k ||C_k psi||^2 /
 * sum_j a_j ||C_j psi||^2, psi <- sqrt(a_k) C_k psi / sqrt(a_k ||C_k psi||^2), a new
 * threshold drawn) and completes the rest of the step from the jumped state, so
 * every trajectory stays on the user's time grid (classic RK4 at the user's dt).
 *
 * Random numbers: SPRNG (the reference's generator) is not available; uniforms
 * in [0, 1) come from xoshiro256** seeded through SplitMix64 with the seed given
 * to use_mcwf_solver (0: a seed from the clock, as make_sprng_seed does).  Results
 * agree with the reference statistically, not trajectory by trajectory. */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>

#include "quac_internal.h"

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void quac_rng_seed_state(unsigned long long *q, uint64_t seed) {
  if (seed == 0) seed = (uint64_t)time(NULL) ^ 0x5DEECE66DULL;
  uint64_t z = seed;
  for (int i = 0; i < 4; i++) {
    z += 0x9E3779B97F4A7C15ULL;
    uint64_t t = z;
    t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ULL;
    t = (t ^ (t >> 27)) * 0x94D049BB133111EBULL;
    q[i] = t ^ (t >> 31);
  }
}

void quac_rng_seed(qsystem s, uint64_t seed) { quac_rng_seed_state(s->mcwf_rng, seed); }

double quac_rng_uniform(qsystem s) { return quac_rng_uniform_state(s->mcwf_rng); }

double quac_rng_uniform_state(unsigned long long *q) {
  const uint64_t r = rotl64(q[1] * 5, 7) * 9;
  const uint64_t t = q[1] << 17;
  q[2] ^= q[0];
  q[3] ^= q[1];
  q[1] ^= q[2];
  q[0] ^= q[3];
  q[2] ^= t;
  q[3] = rotl64(q[3], 45);
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

/* jump operators of the time-independent Lindblad terms (qsystem.c:393-399) */
static int build_jumps(qsystem s) {
  if (s->mcwf_jump) return 0;
  PetscInt n = 0;
  for (PetscInt i = 0; i < s->num_time_indep; i++)
    if (s->time_indep[i].my_term_type == LINDBLAD) n++;
  if (n == 0) return quac_fail(9, "use_mcwf_solver needs at least one time-independent Lindblad term");
  s->mcwf_jump = (Mat *)calloc((size_t)n, sizeof(Mat));
  s->mcwf_jump_a = (PetscScalar *)calloc((size_t)n, sizeof(PetscScalar));
  s->mcwf_njump = 0;
  for (PetscInt i = 0; i < s->num_time_indep; i++) {
    mat_term *t = &s->time_indep[i];
    if (t->my_term_type != LINDBLAD) continue;
    Mat m = NULL;
    construct_op_matrix_wf_list(s, 1.0, &m, t->num_ops, t->ops);
    if (!m) return 9;
    s->mcwf_jump[s->mcwf_njump] = m;
    s->mcwf_jump_a[s->mcwf_njump] = t->a;
    s->mcwf_njump++;
  }
  return 0;
}

void quac_mcwf_free(qsystem s) {
  for (PetscInt j = 0; j < s->mcwf_njump; j++) qh_mat_destroy(s->mcwf_jump[j]);
  free(s->mcwf_jump);
  free(s->mcwf_jump_a);
  free(s->rand_number);
  s->mcwf_jump = NULL;
  s->mcwf_jump_a = NULL;
  s->rand_number = NULL;
  s->mcwf_njump = 0;
}

/* copy the initial condition into every trajectory and draw the thresholds
 * (qsystem.c:977-989) */
static int spawn(qsystem s, qvec x) {
  if (x->ens_spawned) {
    /* trajectories written one by one (add_to_wf_ens_loc): only the thresholds */
    if (x->n_ensemble != s->num_local_trajs)
      return quac_fail(9, "the ensemble holds %lld trajectories, the system %lld", (long long)x->n_ensemble,
                       (long long)s->num_local_trajs);
    if (!s->rand_number) {
      s->rand_number = (PetscReal *)malloc(sizeof(PetscReal) * (size_t)x->n_ensemble);
      for (int64_t i = 0; i < x->n_ensemble; i++) s->rand_number[i] = quac_rng_uniform(s);
    }
    return 0;
  }
  const int64_t M = s->num_local_trajs, N = x->n;
  Vec block = NULL;
  if (qh_vec_create(M * N, QH_FP64, &block)) return quac_fail(9, "cannot allocate %lld trajectories: %s",
                                                              (long long)M, qh_last_error());
  if (qh_ens_broadcast(x->data, block, M)) {
    qh_vec_destroy(block);
    return quac_fail(9, "ensemble set-up failed: %s", qh_last_error());
  }
  x->ens_datas = (Vec *)malloc(sizeof(Vec));
  x->ens_datas[0] = block;
  x->n_ensemble = M;
  free(s->rand_number);
  s->rand_number = (PetscReal *)malloc(sizeof(PetscReal) * (size_t)M);
  for (int64_t i = 0; i < M; i++) s->rand_number[i] = quac_rng_uniform(s);
  x->ens_spawned = PETSC_TRUE;
  return 0;
}

/* choose the jump operator from the squared norms n2[j] = ||C_j psi||^2
 * (qsystem.c:1166-1218): p_j = Re(a_j n2_j), operator k with probability
 * p_k / sum p, scale sqrt(a_k) / sqrt(p_k).  Returns 0 when no operator can
 * act (sum p = 0; no draw is made); p is scratch of mcwf_njump entries. */
static int choose_jump(qsystem s, const double *n2, double *p, int32_t *kout, double complex *sc) {
  double tot = 0.0;
  for (PetscInt j = 0; j < s->mcwf_njump; j++) {
    p[j] = creal(s->mcwf_jump_a[j] * n2[j]);
    tot += p[j];
  }
  if (!(tot > 0.0)) return 0;
  const double u = quac_rng_uniform(s);
  double acc = 0.0;
  PetscInt k = s->mcwf_njump - 1;
  for (PetscInt j = 0; j < s->mcwf_njump; j++) {
    acc += p[j] / tot;
    if (acc >= u) {
      k = j;
      break;
    }
  }
  *kout = (int32_t)k;
  *sc = csqrt(s->mcwf_jump_a[k]) / sqrt(p[k]);
  return 1;
}

/* the jump of trajectory i (qsystem.c:1159-1235) */
static int jump(qsystem s, Vec block, int64_t i, int64_t N, Vec work) {
  Vec v = NULL;
  if (qh_vec_view(block, i * N, N, &v)) return quac_fail(9, "%s", qh_last_error());
  double *n2 = (double *)malloc(sizeof(double) * 2 * (size_t)s->mcwf_njump);
  int rc = 0;
  for (PetscInt j = 0; j < s->mcwf_njump && !rc; j++) {
    n2[j] = 0.0;
    rc = qh_spmv(s->mcwf_jump[j], v, work);
    if (!rc) rc = qh_ens_norms(work, 1, &n2[j]);
  }
  int32_t k = 0;
  double complex sc = 0.0;
  if (!rc && choose_jump(s, n2, n2 + s->mcwf_njump, &k, &sc)) {
    rc = qh_spmv(s->mcwf_jump[k], v, work);
    if (!rc) rc = qh_vec_scale(work, creal(sc), cimag(sc));
    if (!rc) rc = qh_vec_copy(work, v);
  }
  s->rand_number[i] = quac_rng_uniform(s);
  free(n2);
  qh_vec_destroy(v);
  if (rc) return quac_fail(9, "MCWF jump failed: %s", qh_last_error());
  return 0;
}

/* the device-resident loop applies (at most 1024 amplitudes, rows of the
 * generator with at most 16 entries, at most 16 jump operators) and is not
 * switched off by QUAC_MCWF_STEPWISE */
static int fast_path(qsystem s, int64_t N) {
  (void)N;
  if (getenv("QUAC_MCWF_STEPWISE")) return 0;
  int fits = 0;
  if (qh_mcwf_fits(s->mat_A, (int)s->mcwf_njump, &fits)) return 0;
  return fits;
}

/* Trajectories of at most 1024 amplitudes without a monitor: each trajectory runs
 * on the device from one jump to the next (qh_mcwf_advance, which stops it at
 * the located crossing inside its step) and the jumps due are applied together
 * (qh_mcwf_jump_norms, choose_jump, qh_mcwf_jump_apply), until every trajectory
 * has taken all J steps.  The steps, the jump test and
 * the jump are those of the step-by-step loop below; the random numbers of the
 * jumps are drawn round by round in trajectory order, so a given seed yields
 * other (equally distributed) trajectories than the step-by-step loop. */
static int run_to_end(qsystem s, Vec block, int64_t M, int64_t N, int32_t J, double dt, double hlast) {
  const int nj = (int)s->mcwf_njump;
  int32_t *jst = (int32_t *)calloc((size_t)M, sizeof(int32_t));
  double *tsub = (double *)calloc((size_t)M, sizeof(double));
  int32_t *jmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)M);
  int32_t *kk = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
  double *sre = (double *)malloc(sizeof(double) * (size_t)M);
  double *sim = (double *)malloc(sizeof(double) * (size_t)M);
  double *n2 = (double *)malloc(sizeof(double) * (size_t)M * (size_t)nj);
  double *p = (double *)malloc(sizeof(double) * (size_t)nj);
  int rc = (jst && tsub && jmp && idx && kk && sre && sim && n2 && p) ? 0 : quac_fail(9, "out of host memory");
  while (!rc) {
    if (qh_mcwf_advance(s->mat_A, block, M, s->rand_number, jst, tsub, J, dt, hlast, jmp)) {
      rc = quac_fail(9, "MCWF steps failed: %s", qh_last_error());
      break;
    }
    int32_t n = 0;
    for (int64_t i = 0; i < M; i++)
      if (jmp[i]) idx[n++] = i;
    if (n == 0) break;
    if (qh_mcwf_jump_norms(nj, s->mcwf_jump, block, N, idx, n, n2)) {
      rc = quac_fail(9, "MCWF jump failed: %s", qh_last_error());
      break;
    }
    /* draws in the order of jump(): the operator, then the next threshold;
     * the trajectories that jump are compacted to the front of idx */
    int32_t m = 0;
    for (int32_t q = 0; q < n; q++) {
      const int64_t i = idx[q];
      double complex sc = 0.0;
      if (choose_jump(s, n2 + (size_t)q * nj, p, &kk[m], &sc)) {
        idx[m] = i;
        sre[m] = creal(sc);
        sim[m] = cimag(sc);
        m++;
      }
      s->rand_number[i] = quac_rng_uniform(s);
    }
    if (m && qh_mcwf_jump_apply(nj, s->mcwf_jump, block, N, idx, m, kk, sre, sim))
      rc = quac_fail(9, "MCWF jump failed: %s", qh_last_error());
  }
  free(jst);
  free(tsub);
  free(jmp);
  free(idx);
  free(kk);
  free(sre);
  free(sim);
  free(n2);
  free(p);
  return rc;
}

/* trajectory i fell below its threshold during a step of size h that started
 * from its state in pre: locate the crossing (qh_mcwf_cross), jump there, and
 * take the rest of the step from the jumped state, again locating a crossing
 * of the new threshold if one falls inside it */
static int jump_in_step(qsystem s, Vec pre, Vec block, int64_t i, int64_t N, double h, Vec work) {
  Vec p = NULL, c = NULL;
  if (qh_vec_view(pre, i * N, N, &p) || qh_vec_view(block, i * N, N, &c)) {
    qh_vec_destroy(p);
    return quac_fail(9, "%s", qh_last_error());
  }
  double rem = h;
  int rc = 0;
  for (int pass = 0; pass < 1000 && !rc; pass++) {
    double tau = rem;
    if (qh_mcwf_cross(s->mat_A, p, c, rem, s->rand_number[i], &tau)) {
      rc = quac_fail(9, "MCWF crossing failed: %s", qh_last_error());
      break;
    }
    rc = jump(s, block, i, N, work);
    rem -= tau;
    if (rc || rem <= 1e-14 * h) break;
    double n2 = 0.0;
    if (qh_vec_copy(c, p) || qh_ens_rk4(s->mat_A, c, 1, rem, 1) || qh_ens_norms(c, 1, &n2)) {
      rc = quac_fail(9, "MCWF step failed: %s", qh_last_error());
      break;
    }
    if (n2 >= s->rand_number[i]) break;
  }
  qh_vec_destroy(p);
  qh_vec_destroy(c);
  return rc;
}

int quac_mcwf_time_step(qsystem s, qvec x, PetscReal t0, PetscReal tmax, PetscReal dt, PetscInt steps_max,
                        int (*monitor)(qsystem, PetscInt, PetscReal, qvec), PetscInt *nsteps, PetscInt step0,
                        int call_initial) {
  if (x->my_type != WF_ENSEMBLE) return quac_fail(0, "time_step_sys with use_mcwf_solver needs a qvec from create_qvec_sys");
  if (build_jumps(s) || spawn(s, x)) return 9;
  const int64_t M = x->n_ensemble, N = x->n;
  Vec block = x->ens_datas[0];
  Vec work = NULL;
  if (qh_vec_create(N, QH_FP64, &work)) return quac_fail(9, "%s", qh_last_error());
  double *nrm = (double *)malloc(sizeof(double) * (size_t)M);
  PetscReal t = t0;
  PetscInt step = 0;
  const PetscReal eps = 1e-12 * (fabs(tmax) + fabs(dt) + 1.0);
  int rc = 0;
  if (monitor && call_initial) rc = monitor(s, step0, t, x);
  if (!monitor && fast_path(s, N)) {
    /* the step sizes of the loop below: dt, the last one cut at tmax */
    PetscReal hlast = dt;
    while (t < tmax - eps && step < steps_max && step < INT32_MAX) {
      hlast = (t + dt > tmax - eps) ? tmax - t : dt;
      t += hlast;
      step++;
    }
    if (step > 0) rc = run_to_end(s, block, M, N, (int32_t)step, dt, hlast);
  }
  /* step by step: every trajectory takes the step; the states at the step's
   * start are kept to locate the crossings of those that fell below threshold */
  Vec pre = NULL;
  if (!rc && t < tmax - eps && step < steps_max && qh_vec_create(M * N, QH_FP64, &pre))
    rc = quac_fail(9, "%s", qh_last_error());
  while (!rc && t < tmax - eps && step < steps_max) {
    PetscReal h = dt;
    if (t + h > tmax - eps) h = tmax - t;
    if (qh_vec_copy(block, pre) || qh_ens_rk4(s->mat_A, block, M, h, 1) || qh_ens_norms(block, M, nrm)) {
      rc = quac_fail(9, "MCWF step failed: %s", qh_last_error());
      break;
    }
    for (int64_t i = 0; i < M && !rc; i++)
      if (nrm[i] < s->rand_number[i]) rc = jump_in_step(s, pre, block, i, N, h, work);
    t += h;
    step++;
    if (!rc && monitor) rc = monitor(s, step0 + step, t, x);
  }
  qh_vec_destroy(pre);
  free(nrm);
  qh_vec_destroy(work);
  s->current_time = t;
  if (nsteps) *nsteps = step;
  return rc;
}

/* ensemble average of <psi_i|O|psi_i> / <psi_i|psi_i> (qvec_utilities.c:1353-1420) */
int quac_mcwf_expect(qvec x, qc_chain *c, double *re, double *im) {
  return qh_ens_expect(x->ens_datas[0], x->n_ensemble, x->n, c->nops, c->ops, c->nlev, c->lev, re, im);
}
