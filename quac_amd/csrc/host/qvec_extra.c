/* qvec_extra.c -- state comparisons and reductions of qvec_utilities.c that lie
 * beside the stepping path: |psi><psi|, partial traces, fidelity, superfidelity,
 * Hilbert-Schmidt distance, equality, text printers and the (op, fock) location
 * helper.  They copy the state(s) to the host and work there (like the
 * reference, which gathers the state to rank 0 for most of them).
 *
 * Density-matrix element (row, col) is entry N col + row of the vector; in a
 * composite index subsystem 0 is the most significant digit. */
#include <complex.h>
#include <math.h>
#include <string.h>

#include "quac_internal.h"

/* the state as host complex128 (caller frees) */
static double complex *host_state(qvec q) {
  double complex *h = (double complex *)malloc(sizeof(double complex) * (size_t)q->n);
  if (!h) {
    quac_fail(9, "out of host memory");
    return NULL;
  }
  if (qh_vec_get(q->data, (double *)h)) {
    free(h);
    quac_fail(9, "state copy failed: %s", qh_last_error());
    return NULL;
  }
  return h;
}

/* qvec_utilities.c:57-90: destination (N^2) = |psi><psi| of source (N) */
void copy_qvec_wf_to_dm(qvec source, qvec destination) {
  if (source->n * source->n != destination->n) QUAC_FAIL(1, "wf and dm do not have consistent sizes!");
  double complex *w = host_state(source);
  if (!w) return;
  const PetscInt n = source->n;
  double complex *d = (double complex *)malloc(sizeof(double complex) * (size_t)(n * n));
  if (!d) {
    free(w);
    QUAC_FAIL(9, "out of host memory");
  }
  for (PetscInt c = 0; c < n; c++)
    for (PetscInt r = 0; r < n; r++) d[n * c + r] = w[r] * conj(w[c]);
  const int rc = qh_vec_set(destination->data, (const double *)d);
  free(w);
  free(d);
  if (rc) QUAC_FAIL(9, "copy_qvec_wf_to_dm: %s", qh_last_error());
}

/* qvec_utilities.c:26-30 (VecEqual: every entry identical) */
void check_qvec_equal(qvec state1, qvec state2, PetscBool *flag) {
  *flag = PETSC_FALSE;
  if (state1->n != state2->n || state1->my_type != state2->my_type) return;
  double complex *a = host_state(state1), *b = a ? host_state(state2) : NULL;
  if (a && b) *flag = memcmp(a, b, sizeof(double complex) * (size_t)state1->n) == 0 ? PETSC_TRUE : PETSC_FALSE;
  free(a);
  free(b);
}

/* qvec_utilities.c:1591-1656 with 1441-1590: the reduced density matrix over
 * the subsystems not listed, kept in their original order (the first kept
 * subsystem the most significant): rho_red(r, c) = sum_t rho((r, t), (c, t)),
 * rho = |psi><psi| for a wavefunction.  The reference's density-matrix branch
 * writes element (r, c) at (c, r) and makes the first kept subsystem the least
 * significant; its tests use real symmetric states kept to one subsystem,
 * where the two agree. */
void ptrace_over_list_qvec(qvec full_state, PetscInt n_ops, PetscInt *op_loc_list, qvec *new_dm) {
  const int dm = full_state->my_type == DENSITY_MATRIX;
  if (!dm && full_state->my_type != WAVEFUNCTION)
    QUAC_FAIL(9, "ptrace_over_list_qvec: a wavefunction or a density matrix is required");
  const PetscInt nsub = full_state->n_ops;
  if (n_ops > nsub) QUAC_FAIL(9, "Trying to ptrace more ops than are in the state!");
  int *traced = (int *)calloc((size_t)nsub, sizeof(int));
  for (PetscInt i = 0; i < n_ops; i++) {
    if (op_loc_list[i] < 0 || op_loc_list[i] > nsub - 1) {
      free(traced);
      QUAC_FAIL(9, "op_loc_list has an index greater than nops in state!");
    }
    traced[op_loc_list[i]] = 1;
  }
  PetscInt nkeep = 0, Nk = 1, Nt = 1;
  for (PetscInt i = 0; i < nsub; i++) {
    if (traced[i]) Nt *= full_state->hspace_dims[i];
    else {
      Nk *= full_state->hspace_dims[i];
      nkeep++;
    }
  }
  if (nkeep == 0) {
    free(traced);
    QUAC_FAIL(9, "ptrace_over_list_qvec: every subsystem is traced over");
  }
  PetscInt *dims_new = (PetscInt *)malloc(sizeof(PetscInt) * (size_t)(2 * nkeep));
  for (PetscInt i = 0, k = 0; i < nsub; i++)
    if (!traced[i]) {
      dims_new[k] = dims_new[k + nkeep] = full_state->hspace_dims[i];
      k++;
    }
  /* full index of the (kept digits kk, traced digits tt) pair */
  const PetscInt N = Nk * Nt;
  PetscInt *full_of = (PetscInt *)malloc(sizeof(PetscInt) * (size_t)N);
  for (PetscInt kk = 0; kk < Nk; kk++)
    for (PetscInt tt = 0; tt < Nt; tt++) {
      PetscInt rk = kk, rt = tt, idx = 0, mult = 1;
      for (PetscInt i = nsub - 1; i >= 0; i--) { /* least significant subsystem first */
        const PetscInt L = full_state->hspace_dims[i];
        PetscInt digit;
        if (traced[i]) {
          digit = rt % L;
          rt /= L;
        } else {
          digit = rk % L;
          rk /= L;
        }
        idx += digit * mult;
        mult *= L;
      }
      full_of[kk * Nt + tt] = idx;
    }
  double complex *x = host_state(full_state);
  double complex *red = (double complex *)calloc((size_t)(Nk * Nk), sizeof(double complex));
  if (x && red) {
    for (PetscInt c = 0; c < Nk; c++)
      for (PetscInt r = 0; r < Nk; r++) {
        double complex acc = 0.0;
        for (PetscInt t = 0; t < Nt; t++) {
          const PetscInt fr = full_of[r * Nt + t], fc = full_of[c * Nt + t];
          acc += dm ? x[N * fc + fr] : x[fr] * conj(x[fc]);
        }
        red[Nk * c + r] = acc;
      }
    create_arb_qvec_dims(new_dm, 2 * nkeep, dims_new, DENSITY_MATRIX);
    if (*new_dm && qh_vec_set((*new_dm)->data, (const double *)red))
      quac_fail(9, "ptrace_over_list_qvec: %s", qh_last_error());
  }
  free(x);
  free(red);
  free(full_of);
  free(dims_new);
  free(traced);
}

/* eigenvalues (w) and, if V, eigenvectors (columns of V, row-major n x n) of a
 * real symmetric row-major matrix A (destroyed): cyclic Jacobi rotations
 * (Golub & Van Loan, Matrix Computations, sec. 8.5) until the off-diagonal
 * mass is below 1e-30 of the total; above n = 128 the device eigensolver
 * (qh_sym_eig, O(n^3) once instead of per sweep) */
static void sym_eig_jacobi(int n, double *A, double *w, double *V);
static int sym_eig(int n, double *A, double *w, double *V) {
  if (n <= 128) {
    sym_eig_jacobi(n, A, w, V);
    return 0;
  }
  double *C = V ? (double *)malloc(sizeof(double) * (size_t)n * (size_t)n) : NULL;
  if (V && !C) return quac_fail(9, "out of host memory"), -1;
  if (qh_sym_eig(n, A, w, C)) {
    free(C);
    return quac_fail(9, "fidelity eigensolver: %s", qh_last_error()), -1;
  }
  if (V) /* column-major C (column k = eigenvector k) to row-major V */
    for (int i = 0; i < n; i++)
      for (int k = 0; k < n; k++) V[(size_t)i * n + k] = C[(size_t)k * n + i];
  free(C);
  return 0;
}

static void sym_eig_jacobi(int n, double *A, double *w, double *V) {
  if (V)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) V[i * n + j] = i == j;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p * n + q];
        if (apq == 0.0) continue;
        const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) { /* columns p, q */
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) { /* rows p, q */
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        if (V)
          for (int k = 0; k < n; k++) {
            const double vkp = V[k * n + p], vkq = V[k * n + q];
            V[k * n + p] = c * vkp - s * vkq;
            V[k * n + q] = s * vkp + c * vkq;
          }
      }
  }
  for (int i = 0; i < n; i++) w[i] = A[i * n + i];
}

/* real 2N x 2N embedding [[Re M, -Im M], [Im M, Re M]] (row-major) of the
 * N x N matrix stored column-major in x; the embedding commutes with products,
 * adjoints and functions of Hermitian matrices, and a Hermitian M's embedding
 * has each eigenvalue of M twice */
static void embed(PetscInt N, const double complex *x, double *E) {
  const PetscInt M = 2 * N;
  for (PetscInt r = 0; r < N; r++)
    for (PetscInt c = 0; c < N; c++) {
      const double complex v = x[N * c + r];
      E[r * M + c] = creal(v);
      E[r * M + c + N] = -cimag(v);
      E[(r + N) * M + c] = cimag(v);
      E[(r + N) * M + c + N] = creal(v);
    }
}

static void matmul(PetscInt n, const double *A, const double *B, double *C) {
  for (PetscInt i = 0; i < n; i++) {
    for (PetscInt j = 0; j < n; j++) C[i * n + j] = 0.0;
    for (PetscInt k = 0; k < n; k++) {
      const double a = A[i * n + k];
      if (a == 0.0) continue;
      for (PetscInt j = 0; j < n; j++) C[i * n + j] += a * B[k * n + j];
    }
  }
}

/* qvec_utilities.c:2339-2430: (Tr sqrt(sqrt(rho) sigma sqrt(rho)))^2, negative
 * eigenvalues dropped */
static int fidelity_dm_dm(PetscInt N, const double complex *rho, const double complex *sig, double *fid) {
  const PetscInt M = 2 * N;
  double *R = (double *)malloc(sizeof(double) * (size_t)(M * M)), *V = (double *)malloc(sizeof(double) * (size_t)(M * M));
  double *S = (double *)malloc(sizeof(double) * (size_t)(M * M)), *T = (double *)malloc(sizeof(double) * (size_t)(M * M));
  double *w = (double *)malloc(sizeof(double) * (size_t)M);
  if (!R || !V || !S || !T || !w) {
    free(R), free(V), free(S), free(T), free(w);
    return quac_fail(9, "out of host memory"), -1;
  }
  embed(N, rho, R);
  if (sym_eig((int)M, R, w, V)) {
    free(R), free(V), free(S), free(T), free(w);
    return -1;
  }
  /* R = sqrt(rho) embedded = V diag(sqrt(max(w, 0))) V^T */
  for (PetscInt i = 0; i < M; i++)
    for (PetscInt k = 0; k < M; k++) T[i * M + k] = V[i * M + k] * sqrt(w[k] > 0 ? w[k] : 0.0);
  for (PetscInt i = 0; i < M; i++)
    for (PetscInt j = 0; j < M; j++) {
      double acc = 0.0;
      for (PetscInt k = 0; k < M; k++) acc += T[i * M + k] * V[j * M + k];
      R[i * M + j] = acc;
    }
  embed(N, sig, S);
  matmul(M, R, S, T);
  matmul(M, T, R, S);
  /* symmetrise against rounding before the eigenvalues */
  for (PetscInt i = 0; i < M; i++)
    for (PetscInt j = i + 1; j < M; j++) S[i * M + j] = S[j * M + i] = 0.5 * (S[i * M + j] + S[j * M + i]);
  if (sym_eig((int)M, S, w, NULL)) {
    free(R), free(V), free(S), free(T), free(w);
    return -1;
  }
  double f = 0.0;
  for (PetscInt i = 0; i < M; i++)
    if (w[i] > 0) f += sqrt(w[i]);
  f *= 0.5; /* every eigenvalue appears twice in the embedding */
  *fid = f * f;
  free(R), free(V), free(S), free(T), free(w);
  return 0;
}

/* <psi|rho|psi> (qvec_utilities.c:2271-2337) */
static double fidelity_dm_wf(PetscInt N, const double complex *rho, const double complex *psi) {
  double complex acc = 0.0;
  for (PetscInt c = 0; c < N; c++) {
    double complex col = 0.0;
    for (PetscInt r = 0; r < N; r++) col += conj(psi[r]) * rho[N * c + r];
    acc += col * psi[c];
  }
  return creal(acc);
}

/* |<b|a>|^2 (qvec_utilities.c:2247-2252) */
static double fidelity_wf_wf(PetscInt n, const double complex *a, const double complex *b) {
  double complex acc = 0.0;
  for (PetscInt i = 0; i < n; i++) acc += a[i] * conj(b[i]);
  return creal(acc) * creal(acc) + cimag(acc) * cimag(acc);
}

/* qvec_utilities.c:2159-2245: fidelity of two states; for a spawned MCWF
 * ensemble against a wavefunction the mean over the trajectories of
 * |<psi|traj_i>|^2 / <traj_i|traj_i>, and in *fid_var the sample variance of
 * the unnormalised |<psi|traj_i>|^2 about that mean (the reference's form) */
void get_fidelity_qvec(qvec q1, qvec q2, PetscReal *fidelity, PetscReal *fid_var) {
  *fid_var = 0.0;
  *fidelity = 0.0;
  const qvec_type t1 = q1->my_type, t2 = q2->my_type;
  if ((t1 == WF_ENSEMBLE && t2 == WAVEFUNCTION) || (t2 == WF_ENSEMBLE && t1 == WAVEFUNCTION)) {
    qvec ens = t1 == WF_ENSEMBLE ? q1 : q2, wf = t1 == WF_ENSEMBLE ? q2 : q1;
    double complex *b = host_state(wf);
    if (!b) return;
    if (!ens->ens_spawned) {
      double complex *a = host_state(ens);
      if (a) *fidelity = fidelity_wf_wf(wf->n, a, b);
      free(a);
      free(b);
      return;
    }
    const PetscInt M = ens->n_ensemble, N = ens->n;
    double complex *a = (double complex *)malloc(sizeof(double complex) * (size_t)(M * N));
    double *f = (double *)malloc(sizeof(double) * (size_t)M);
    if (a && f && !qh_vec_get(ens->ens_datas[0], (double *)a)) {
      double mean = 0.0;
      for (PetscInt i = 0; i < M; i++) {
        double nrm = 0.0;
        for (PetscInt k = 0; k < N; k++) nrm += creal(a[i * N + k] * conj(a[i * N + k]));
        f[i] = fidelity_wf_wf(N, a + i * N, b);
        mean += f[i] / nrm;
      }
      mean /= (double)M;
      double var = 0.0;
      for (PetscInt i = 0; i < M; i++) var += (f[i] - mean) * (f[i] - mean);
      *fidelity = mean;
      *fid_var = M > 1 ? var / (double)(M - 1) : 0.0;
    } else {
      quac_fail(9, "get_fidelity_qvec: ensemble copy failed");
    }
    free(a);
    free(f);
    free(b);
    return;
  }
  double complex *a = host_state(q1), *b = a ? host_state(q2) : NULL;
  if (!a || !b) {
    free(a);
    return;
  }
  if (t1 == WAVEFUNCTION && t2 == WAVEFUNCTION) {
    if (q1->n != q2->n) quac_fail(0, "The wavefunctions are not the same size! Fidelity cannot be calculated.");
    else *fidelity = fidelity_wf_wf(q1->n, a, b);
  } else if (t1 == DENSITY_MATRIX && t2 == DENSITY_MATRIX) {
    if (q1->n != q2->n) quac_fail(0, "The input density matrices are not the same size! Fidelity cannot be calculated.");
    else fidelity_dm_dm(q1->total_levels, a, b, fidelity);
  } else if ((t1 == DENSITY_MATRIX && t2 == WAVEFUNCTION) || (t1 == WAVEFUNCTION && t2 == DENSITY_MATRIX)) {
    qvec d = t1 == DENSITY_MATRIX ? q1 : q2, w = t1 == DENSITY_MATRIX ? q2 : q1;
    if (d->n != w->n * w->n)
      quac_fail(0, "The input density matrix and wavefunction are not compatible! Fidelity cannot be calculated.");
    else *fidelity = fidelity_dm_wf(w->n, t1 == DENSITY_MATRIX ? a : b, t1 == DENSITY_MATRIX ? b : a);
  } else {
    quac_fail(9, "Types not understood in get_fidelity_qvec!");
  }
  free(a);
  free(b);
}

/* PETSc VecDot(x, y) = sum x_i conj(y_i) */
static double complex vdot(PetscInt n, const double complex *x, const double complex *y) {
  double complex acc = 0.0;
  for (PetscInt i = 0; i < n; i++) acc += x[i] * conj(y[i]);
  return acc;
}

/* qvec_utilities.c:2102-2141: sqrt(Re(Tr(rho1 rho2) + sqrt(1 - Tr rho1^2) sqrt(1 - Tr rho2^2)));
 * a wavefunction argument is first made |psi><psi| */
void get_superfidelity_qvec(qvec q1, qvec q2, PetscReal *superfidelity) {
  *superfidelity = 0.0;
  if (q1->my_type == WAVEFUNCTION && q2->my_type == WAVEFUNCTION)
    QUAC_FAIL(9, "Calculating superfidelity between two wavefunctions is not recommended. Use get_superfidelity_qvec!");
  if ((q1->my_type != DENSITY_MATRIX && q1->my_type != WAVEFUNCTION) ||
      (q2->my_type != DENSITY_MATRIX && q2->my_type != WAVEFUNCTION))
    QUAC_FAIL(9, "Types not understand in get_superfidelity_qvec!");
  qvec d1 = q1, d2 = q2, tmp = NULL;
  if (q1->my_type == WAVEFUNCTION || q2->my_type == WAVEFUNCTION) {
    qvec w = q1->my_type == WAVEFUNCTION ? q1 : q2, d = q1->my_type == WAVEFUNCTION ? q2 : q1;
    create_arb_qvec(&tmp, d->n, DENSITY_MATRIX);
    if (!tmp) return;
    copy_qvec_wf_to_dm(w, tmp);
    d1 = d;
    d2 = tmp;
  }
  if (d1->n != d2->n) {
    destroy_qvec(&tmp);
    QUAC_FAIL(9, "get_superfidelity_qvec: sizes differ");
  }
  double complex *a = host_state(d1), *b = a ? host_state(d2) : NULL;
  if (a && b) {
    const double complex v1 = vdot(d1->n, a, b), v2 = vdot(d1->n, a, a), v3 = vdot(d1->n, b, b);
    *superfidelity = sqrt(creal(v1 + csqrt(1.0 - v2) * csqrt(1.0 - v3)));
  }
  free(a);
  free(b);
  destroy_qvec(&tmp);
}

/* qvec_utilities.c:1657-1680: ||rho1 - rho2||_F^2 */
void get_hilbert_schmidt_dist_qvec(qvec q1, qvec q2, PetscReal *hs_dist) {
  *hs_dist = 0.0;
  if (q1->my_type != DENSITY_MATRIX || q2->my_type != DENSITY_MATRIX) QUAC_FAIL(9, "HS is only for DM-DM comparisons");
  if (q1->n != q2->n) QUAC_FAIL(9, "DMs are not the same size in hs_dist");
  double complex *a = host_state(q1), *b = a ? host_state(q2) : NULL;
  if (a && b) {
    double acc = 0.0;
    for (PetscInt i = 0; i < q1->n; i++) {
      const double complex d = a[i] - b[i];
      acc += creal(d) * creal(d) + cimag(d) * cimag(d);
    }
    *hs_dist = acc;
  }
  free(a);
  free(b);
}

/* qvec_utilities.c:291-305 */
void print_dm_qvec(qvec dm) {
  double complex *x = host_state(dm);
  if (!x) return;
  const PetscInt N = dm->total_levels;
  for (PetscInt i = 0; i < N; i++) {
    for (PetscInt j = 0; j < N; j++) printf("%4.3e + %4.3ei ", creal(x[N * j + i]), cimag(x[N * j + i]));
    printf("\n");
  }
  printf("\n");
  free(x);
}

/* qvec_utilities.c:397-412 */
void print_wf_qvec(qvec state) {
  double complex *x = host_state(state);
  if (!x) return;
  for (PetscInt i = 0; i < state->n; i++) printf("%4.3e + %4.3ei\n", creal(x[i]), cimag(x[i]));
  printf("\n");
  free(x);
}

/* qvec_utilities.c:971-993: (operator, fock state) pairs */
void get_qvec_loc_fock_op(qvec state, PetscInt *loc, PetscInt num_ops, ...) {
  if (num_ops < 1 || num_ops > 64) QUAC_FAIL(9, "get_qvec_loc_fock_op: 1..64 operators");
  operator ops[64];
  PetscInt fock[64];
  va_list ap;
  va_start(ap, num_ops);
  for (PetscInt i = 0; i < num_ops; i++) {
    ops[i] = va_arg(ap, operator);
    fock[i] = va_arg(ap, PetscInt);
  }
  va_end(ap);
  get_qvec_loc_fock_op_list(state, loc, num_ops, ops, fock);
}

/* ------------------------------------------- cross-entropy benchmark scores */
/* Scores of an "experimental" bitstring distribution p_exp (with per-bitstring
 * variances v_exp) against the ideal one p_ref over a Hilbert space of size D
 * (qvec_utilities.c:1720-1991): linear XEB D sum p_exp p_ref - 1, its unbiased
 * form (D sum p_exp p_ref - 1) / (D sum p_ref^2 - 1), log XEB
 * log D + sum p_exp log p_ref + gamma_Euler (bitstrings with p_ref = 0 skipped
 * in the mean and counted with log(3e-33) in the variance), and the heavy-output
 * score sum over p_ref >= ln 2 / D of p_exp with its fidelity (2 h - 1) / ln 2. */
static int same_size(PetscInt a, PetscInt b, const char *fn) {
  if (a == b) return 0;
  quac_fail(9, "Ref and Exp are not the same size in %s!", fn);
  return -1;
}

void get_linear_xeb_fidelity_probs(PetscReal *probs_ref, PetscReal *vars_ref, PetscInt nloc_ref, PetscReal *probs_exp,
                                   PetscReal *vars_exp, PetscInt nloc_exp, PetscInt h_dim, PetscReal *fid,
                                   PetscReal *var) {
  (void)vars_ref;
  if (same_size(nloc_ref, nloc_exp, "get_linear_xeb_fidelity_probs")) return;
  double s = 0.0, v = 0.0;
  for (PetscInt i = 0; i < nloc_ref; i++) {
    s += probs_exp[i] * probs_ref[i];
    v += vars_exp[i] * probs_ref[i] * probs_ref[i];
  }
  *fid = (double)h_dim * s - 1.0;
  *var = (double)h_dim * (double)h_dim * v;
}

void get_linear_xeb_unbiased_fidelity_probs(PetscReal *probs_ref, PetscReal *vars_ref, PetscInt nloc_ref,
                                            PetscReal *probs_exp, PetscReal *vars_exp, PetscInt nloc_exp,
                                            PetscInt h_dim, PetscReal *fid, PetscReal *var) {
  (void)vars_ref;
  if (same_size(nloc_ref, nloc_exp, "get_linear_xeb_unbiased_fidelity_probs")) return;
  double num = 0.0, den = 0.0, v = 0.0;
  for (PetscInt i = 0; i < nloc_ref; i++) {
    num += probs_exp[i] * probs_ref[i];
    den += probs_ref[i] * probs_ref[i];
    v += vars_exp[i] * probs_ref[i] * probs_ref[i];
  }
  const double d = (double)h_dim * den - 1.0;
  *fid = ((double)h_dim * num - 1.0) / d;
  *var = (double)h_dim * (double)h_dim * v / (d * d);
}

void get_log_xeb_fidelity_probs(PetscReal *probs_ref, PetscReal *vars_ref, PetscInt nloc_ref, PetscReal *probs_exp,
                                PetscReal *vars_exp, PetscInt nloc_exp, PetscInt h_dim, PetscReal *fid,
                                PetscReal *var) {
  (void)vars_ref;
  if (same_size(nloc_ref, nloc_exp, "get_log_xeb_fidelity_probs")) return;
  const double euler_gamma = 0.57721566490153286;
  const double log_floor = log(3e-33);
  double s = 0.0, v = 0.0;
  for (PetscInt i = 0; i < nloc_ref; i++) {
    if (probs_ref[i] == 0.0) {
      v += vars_exp[i] * log_floor * log_floor;
    } else {
      const double l = log(probs_ref[i]);
      s += probs_exp[i] * l;
      v += vars_exp[i] * l * l;
    }
  }
  *fid = log((double)h_dim) + s + euler_gamma;
  *var = v;
}

void get_hog_score_probs(PetscReal *probs_ref, PetscReal *vars_ref, PetscInt nloc_ref, PetscReal *probs_exp,
                         PetscReal *vars_exp, PetscInt nloc_exp, PetscInt h_dim, PetscReal *hog_score,
                         PetscReal *hog_var) {
  (void)vars_ref;
  if (same_size(nloc_ref, nloc_exp, "get_hog_score_probs")) return;
  const double heavy = log(2.0) / (double)h_dim;
  double s = 0.0, v = 0.0;
  for (PetscInt i = 0; i < nloc_ref; i++)
    if (probs_ref[i] >= heavy) {
      s += probs_exp[i];
      v += vars_exp[i];
    }
  *hog_score = s;
  *hog_var = v;
}

void get_hog_score_fidelity_probs(PetscReal *probs_ref, PetscReal *vars_ref, PetscInt nloc_ref, PetscReal *probs_exp,
                                  PetscReal *vars_exp, PetscInt nloc_exp, PetscInt h_dim, PetscReal *fid,
                                  PetscReal *var) {
  double h = 0.0, hv = 0.0;
  get_hog_score_probs(probs_ref, vars_ref, nloc_ref, probs_exp, vars_exp, nloc_exp, h_dim, &h, &hv);
  *fid = (2.0 * h - 1.0) / log(2.0);
  *var = (2.0 / log(2.0)) * (2.0 / log(2.0)) * hv;
}

typedef void (*probs_score)(PetscReal *, PetscReal *, PetscInt, PetscReal *, PetscReal *, PetscInt, PetscInt,
                            PetscReal *, PetscReal *);

/* the score of two states' bitstring distributions (diagonals) */
static void score_states(qvec ref, qvec exp, probs_score f, PetscReal *a, PetscReal *b, const char *fn) {
  if (ref->total_levels != exp->total_levels) QUAC_FAIL(9, "Ref and Exp are not the same size in %s!", fn);
  PetscInt nr = 0, ne = 0;
  PetscReal *pr = NULL, *vr = NULL, *pe = NULL, *ve = NULL;
  get_bitstring_probs(ref, &nr, &pr, &vr);
  if (pr) get_bitstring_probs(exp, &ne, &pe, &ve);
  if (pr && pe) f(pr, vr, nr, pe, ve, ne, ref->total_levels, a, b);
  free(pr), free(vr), free(pe), free(ve);
}

void get_linear_xeb_fidelity(qvec ref, qvec exp, PetscReal *fid, PetscReal *var) {
  score_states(ref, exp, get_linear_xeb_fidelity_probs, fid, var, "get_linear_xeb_fidelity");
}
void get_linear_xeb_unbiased_fidelity(qvec ref, qvec exp, PetscReal *fid, PetscReal *var) {
  score_states(ref, exp, get_linear_xeb_unbiased_fidelity_probs, fid, var, "get_linear_xeb_unbiased_fidelity");
}
void get_log_xeb_fidelity(qvec ref, qvec exp, PetscReal *fid, PetscReal *var) {
  score_states(ref, exp, get_log_xeb_fidelity_probs, fid, var, "get_log_xeb_fidelity");
}
void get_hog_score(qvec ref, qvec exp, PetscReal *hog_score, PetscReal *hog_var) {
  score_states(ref, exp, get_hog_score_probs, hog_score, hog_var, "get_hog_score");
}
void get_hog_score_fidelity(qvec ref, qvec exp, PetscReal *fid, PetscReal *var) {
  score_states(ref, exp, get_hog_score_fidelity_probs, fid, var, "get_hog_score_fidelity");
}

/* --------------------------------------------------- single-process accessors */
/* one process holds the whole state, so the _local forms are the global ones
 * (qvec_utilities.c:351-359, 416-449) */
void get_dm_element_qvec_local(qvec dm, PetscInt row, PetscInt col, PetscScalar *val) {
  get_dm_element_qvec(dm, row, col, val);
}
void get_wf_element_qvec_local(qvec wf, PetscInt i, PetscScalar *val) { get_wf_element_qvec(wf, i, val); }

/* amplitude i of trajectory ens_i of a spawned ensemble (the initial state before) */
void get_wf_element_ens_i_qvec_local(qvec wf, PetscInt ens_i, PetscInt i, PetscScalar *val) {
  *val = 0.0;
  if (wf->my_type != WF_ENSEMBLE) QUAC_FAIL(9, "get_wf_element_ens_i_qvec_local: a WF_ENSEMBLE qvec is required");
  if (i < 0 || i >= wf->n) QUAC_FAIL(9, "get_wf_element_ens_i_qvec_local: index out of range");
  if (!wf->ens_spawned) {
    get_wf_element_qvec(wf, i, val);
    return;
  }
  if (ens_i < 0 || ens_i >= wf->n_ensemble) QUAC_FAIL(9, "get_wf_element_ens_i_qvec_local: trajectory out of range");
  double re = 0.0, im = 0.0;
  QUAC_HIPCHK(qh_vec_get_entry(wf->ens_datas[0], ens_i * wf->n + i, &re, &im));
  *val = re + im * I;
}

/* the digits of a global index, the last dimension the least significant
 * (qvec_utilities.c:1427-1438) */
void get_qvec_local_idxs(qvec state, PetscInt global_idx, PetscInt *local_idxs) {
  PetscInt r = global_idx;
  for (PetscInt k = state->ndims_hspace - 1; k >= 0; k--) {
    local_idxs[k] = r % state->hspace_dims[k];
    r /= state->hspace_dims[k];
  }
}

/* qvec_utilities.c:10-24: same type and size */
void check_qvec_consistent(qvec q1, qvec q2) {
  if (q1->my_type != q2->my_type) QUAC_FAIL(9, "qvecs are not the same type!");
  if (q1->n != q2->n) QUAC_FAIL(9, "qvecs are not the same size!");
}

/* qvec_utilities.c:582-626: new subsystem dimensions of the same total size */
void change_qvec_dims(qvec q, PetscInt ndims, PetscInt *dims) {
  PetscInt tot = 1;
  for (PetscInt i = 0; i < ndims; i++) tot *= dims[i];
  if (tot != q->n) QUAC_FAIL(9, "change_qvec_dims: the dimensions do not multiply to the state size");
  PetscInt *nd = (PetscInt *)malloc(sizeof(PetscInt) * (size_t)ndims);
  memcpy(nd, dims, sizeof(PetscInt) * (size_t)ndims);
  free(q->hspace_dims);
  q->hspace_dims = nd;
  q->ndims_hspace = ndims;
  q->n_ops = q->my_type == DENSITY_MATRIX ? ndims / 2 : ndims;
}

/* ------------------------------------------------- classical readout errors */
/* quantum_circuits.c:620-644: mix the bitstring probabilities of qubit
 * `qubit_num` (stride 2^qubit_num): p01 = P(read 0 | state 1), p10 = P(read 1 |
 * state 0), i.e. p0' = (1 - p01) p0 + p10 p1, p1' = p01 p0 + (1 - p10) p1 */
void apply_single_qb_measurement_error_probs(PetscReal *probs, PetscInt n, PetscReal p01, PetscReal p10,
                                             PetscInt qubit_num) {
  const PetscInt n_bef = (PetscInt)1 << qubit_num, n_inc = 2 * n_bef;
  for (PetscInt i = 0; i < n / 2; i++) {
    const PetscInt i0 = (i / n_bef) * n_inc + i % n_bef, i1 = i0 + n_bef;
    const double t0 = probs[i0], t1 = probs[i1];
    probs[i0] = (1 - p01) * t0 + p10 * t1;
    probs[i1] = p01 * t0 + (1 - p10) * t1;
  }
}

/* quantum_circuits.c:646-688: the same mixing applied to the entries of a
 * density matrix's vector, pairs at the stride of the dimensions before
 * `qubit_num` in the state's dimension list */
void apply_single_qb_measurement_error_state(qvec state, PetscReal p01, PetscReal p10, PetscInt qubit_num) {
  if (state->my_type != DENSITY_MATRIX)
    QUAC_FAIL(9, "Must be a DENSITY_MATRIX in apply_single_qb_measurement_error_state");
  if (qubit_num < 0 || qubit_num >= state->ndims_hspace) QUAC_FAIL(9, "qubit out of range");
  PetscInt n_bef = 1;
  for (PetscInt i = 0; i < qubit_num; i++) n_bef *= state->hspace_dims[i];
  const PetscInt n_inc = 2 * n_bef;
  double complex *x = host_state(state);
  if (!x) return;
  for (PetscInt i = 0; i < state->n / 2; i++) {
    const PetscInt i0 = (i / n_bef) * n_inc + i % n_bef, i1 = i0 + n_bef;
    const double complex t0 = x[i0], t1 = x[i1];
    x[i0] = (1 - p01) * t0 + p10 * t1;
    x[i1] = p01 * t0 + (1 - p10) * t1;
  }
  const int rc = qh_vec_set(state->data, (const double *)x);
  free(x);
  if (rc) QUAC_FAIL(9, "apply_single_qb_measurement_error_state: %s", qh_last_error());
}

/* ---------------------------------------------------- projective measurement */
/* the measurement's uniform numbers: xoshiro256** seeded once per process from
 * the quac_seed option (0 or unset: from the clock) */
static unsigned long long g_meas_rng[4];
static int g_meas_seeded = 0;

/* x <- P x on the qubit of op (P: 2 x 2, row-major), for a wavefunction, or
 * P rho P^dag for a density matrix; x is the host copy of the state */
static void project_qubit(qvec q, operator op, const double complex P[2][2], double complex *x) {
  const PetscInt N = q->total_levels, stride = N / (2 * op->n_before);
  if (q->my_type == WAVEFUNCTION) {
    for (PetscInt i = 0; i < N; i++) {
      if ((i / stride) % 2) continue;
      const double complex a0 = x[i], a1 = x[i + stride];
      x[i] = P[0][0] * a0 + P[0][1] * a1;
      x[i + stride] = P[1][0] * a0 + P[1][1] * a1;
    }
    return;
  }
  /* rows (the fast index of the vector), then columns */
  for (PetscInt c = 0; c < N; c++)
    for (PetscInt r = 0; r < N; r++) {
      if ((r / stride) % 2) continue;
      double complex *a0 = &x[N * c + r], *a1 = &x[N * c + r + stride];
      const double complex v0 = *a0, v1 = *a1;
      *a0 = P[0][0] * v0 + P[0][1] * v1;
      *a1 = P[1][0] * v0 + P[1][1] * v1;
    }
  for (PetscInt c = 0; c < N; c++) {
    if ((c / stride) % 2) continue;
    for (PetscInt r = 0; r < N; r++) {
      double complex *b0 = &x[N * c + r], *b1 = &x[N * (c + stride) + r];
      const double complex v0 = *b0, v1 = *b1;
      *b0 = v0 * conj(P[0][0]) + v1 * conj(P[0][1]);
      *b1 = v0 * conj(P[1][0]) + v1 * conj(P[1][1]);
    }
  }
}

/* quantum_circuits.c:484-582: measure each listed Pauli operator (sigma_x / y /
 * z of a qubit) in turn: outcome k (eigenvalue evals[k], eigenvector evecs[k])
 * with probability (1 + evals[k] <op>) / 2, the state projected on it and
 * renormalised; *meas_val is the product of the outcomes' eigenvalues.  (The
 * reference draws eigenvector 0 with probability (1 + <op>) / 2, which is the
 * same for sigma_z, whose evals[0] = +1.) */
void apply_projective_measurement_qubit_list(qvec q, PetscScalar *meas_val, PetscInt num_ops, operator *ops) {
  *meas_val = 1.0;
  if (q->my_type != DENSITY_MATRIX && q->my_type != WAVEFUNCTION)
    QUAC_FAIL(9, "apply_projective_measurement_qubit: a wavefunction or a density matrix is required");
  if (!g_meas_seeded) {
    quac_rng_seed_state(g_meas_rng, (uint64_t)quac_opt_int("quac_seed", 0));
    g_meas_seeded = 1;
  }
  for (PetscInt i = 0; i < num_ops; i++) {
    operator op = ops[i];
    if (op->my_levels != 2 || (op->my_op_type != SIGMA_X && op->my_op_type != SIGMA_Y && op->my_op_type != SIGMA_Z))
      QUAC_FAIL(9, "apply_projective_measurement_qubit: Pauli operators of qubits only");
    PetscScalar ev = 0.0;
    get_expectation_value_qvec_list(q, &ev, 1, &op);
    double p0 = 0.5 * (1.0 + creal(op->evals[0]) * creal(ev));
    p0 = p0 < 0.0 ? 0.0 : p0 > 1.0 ? 1.0 : p0;
    const int k = quac_rng_uniform_state(g_meas_rng) < p0 ? 0 : 1;
    const double prob = k == 0 ? p0 : 1.0 - p0;
    if (!(prob > 0.0)) QUAC_FAIL(9, "apply_projective_measurement_qubit: outcome of zero probability");
    double complex P[2][2];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) P[a][b] = op->evecs[k][a] * conj(op->evecs[k][b]);
    double complex *x = host_state(q);
    if (!x) return;
    project_qubit(q, op, P, x);
    const double sc = q->my_type == DENSITY_MATRIX ? 1.0 / prob : 1.0 / sqrt(prob);
    for (PetscInt j = 0; j < q->n; j++) x[j] *= sc;
    const int rc = qh_vec_set(q->data, (const double *)x);
    free(x);
    if (rc) QUAC_FAIL(9, "apply_projective_measurement_qubit: %s", qh_last_error());
    *meas_val *= op->evals[k];
  }
}

void apply_projective_measurement_qubit(qvec q, PetscScalar *meas_val, PetscInt num_ops, ...) {
  if (num_ops < 1 || num_ops > 64) QUAC_FAIL(9, "apply_projective_measurement_qubit: 1..64 operators");
  operator ops[64];
  va_list ap;
  va_start(ap, num_ops);
  for (PetscInt i = 0; i < num_ops; i++) ops[i] = va_arg(ap, operator);
  va_end(ap);
  apply_projective_measurement_qubit_list(q, meas_val, num_ops, ops);
}

/* ------------------------------------------------- two-qubit concurrence */
/* Wootters' concurrence of a two-qubit density matrix (dm_utilities.c:1487-1602):
 * l1 - l2 - l3 - l4, l_i the square roots, in decreasing order, of the
 * eigenvalues of rho rho~ with rho~ = (sy (x) sy) rho* (sy (x) sy); eigenvalues
 * below 1e-14 count as 0 and, as in dm_utilities.c:1568-1582, the result is not
 * clipped at 0 (Wootters' C is max(0, that)).  The eigenvalues of rho rho~ are
 * those of sqrt(rho) rho~ sqrt(rho), a Hermitian matrix handled on the real
 * embedding as in the fidelity.  rho: 4 x 4 column-major. */
int quac_concurrence_2q(const double complex *rho, double *conc) {
  /* sy (x) sy = [[0,0,0,-1],[0,0,1,0],[0,1,0,0],[-1,0,0,0]] */
  static const int perm[4] = {3, 2, 1, 0};
  static const double sgn[4] = {-1, 1, 1, -1};
  double complex tilde[16];
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++) tilde[4 * c + r] = sgn[r] * sgn[c] * conj(rho[4 * perm[c] + perm[r]]);
  const int M = 8;
  double R[64], V[64], S[64], T[64], w[8];
  embed(4, rho, R);
  sym_eig_jacobi(M, R, w, V);
  for (int i = 0; i < M; i++)
    for (int k = 0; k < M; k++) T[i * M + k] = V[i * M + k] * sqrt(w[k] > 0 ? w[k] : 0.0);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      double acc = 0.0;
      for (int k = 0; k < M; k++) acc += T[i * M + k] * V[j * M + k];
      R[i * M + j] = acc;
    }
  embed(4, tilde, S);
  matmul(M, R, S, T);
  matmul(M, T, R, S);
  for (int i = 0; i < M; i++)
    for (int j = i + 1; j < M; j++) S[i * M + j] = S[j * M + i] = 0.5 * (S[i * M + j] + S[j * M + i]);
  sym_eig_jacobi(M, S, w, NULL);
  /* each eigenvalue twice: sort descending and take every other */
  for (int i = 0; i < M; i++)
    for (int j = i + 1; j < M; j++)
      if (w[j] > w[i]) {
        const double t = w[i];
        w[i] = w[j];
        w[j] = t;
      }
  double l[4];
  for (int i = 0; i < 4; i++) l[i] = w[2 * i] < 1e-14 ? 0.0 : sqrt(w[2 * i]);
  *conc = l[0] - l[1] - l[2] - l[3];
  return 0;
}

/* --------------------------------------------- matrices and ensemble members */
/* qvec_utilities.c:169-195: state <- A state.  For a spawned ensemble every
 * trajectory is multiplied by A.  (The reference's spawned branch multiplies
 * state->data, the initial condition, into every trajectory; here each
 * trajectory's own amplitudes are used, which is what the unspawned and
 * non-ensemble branches do.) */
void qvec_mat_mult(Mat circ_mat, qvec state) {
  int64_t rows = 0, nnz = 0;
  if (qh_mat_info(circ_mat, &rows, &nnz)) QUAC_FAIL(9, "qvec_mat_mult: %s", qh_last_error());
  if (rows != state->n) QUAC_FAIL(9, "qvec_mat_mult: the matrix has %lld rows, the state %lld entries", (long long)rows,
                                  (long long)state->n);
  const int spawned = state->my_type == WF_ENSEMBLE && state->ens_spawned;
  const PetscInt members = spawned ? state->n_ensemble : 1;
  Vec tmp = NULL;
  QUAC_HIPCHK(qh_vec_create(state->n, QH_FP64, &tmp));
  int rc = 0;
  for (PetscInt i = 0; i < members && !rc; i++) {
    Vec v = state->data;
    if (spawned && (rc = qh_vec_view(state->ens_datas[0], i * state->n, state->n, &v))) break;
    rc = qh_spmv(circ_mat, v, tmp);
    if (!rc) rc = qh_vec_copy(tmp, v);
    if (spawned) qh_vec_destroy(v);
  }
  qh_vec_destroy(tmp);
  if (rc) QUAC_FAIL(9, "qvec_mat_mult: %s", qh_last_error());
}

/* trajectories of a WF_ENSEMBLE held in one zeroed device block, marked spawned */
static int ensure_members(qvec state) {
  if (state->ens_spawned) return 0;
  if (state->n_ensemble < 1) return quac_fail(9, "the ensemble has no trajectories"), -1;
  Vec block = NULL;
  if (qh_vec_create(state->n_ensemble * state->n, QH_FP64, &block))
    return quac_fail(9, "cannot allocate the trajectories: %s", qh_last_error()), -1;
  state->ens_datas = (Vec *)malloc(sizeof(Vec));
  state->ens_datas[0] = block;
  state->ens_spawned = PETSC_TRUE;
  return 0;
}

/* qvec_utilities.c:899-911: add alpha to amplitude loc of trajectory i_ens.  The
 * ensemble counts as spawned afterwards: its trajectories start at zero, not
 * at the initial condition, and time stepping does not overwrite them. */
void add_to_wf_ens_loc(qvec state, PetscInt i_ens, PetscScalar alpha, PetscInt loc) {
  if (state->my_type != WF_ENSEMBLE) QUAC_FAIL(9, "add_to_wf_ens_loc: a WF_ENSEMBLE qvec is required");
  if (loc >= state->n || loc < 0) QUAC_FAIL(0, "ERROR! Adding to a qvec location that is out of range!");
  if (i_ens < 0 || i_ens >= state->n_ensemble) QUAC_FAIL(9, "add_to_wf_ens_loc: trajectory out of range");
  if (ensure_members(state)) return;
  QUAC_HIPCHK(qh_vec_add_entry(state->ens_datas[0], i_ens * state->n + loc, creal(alpha), cimag(alpha)));
}

/* qvec_utilities.c:456-466: the amplitudes of trajectory state->ens_i */
void print_wf_ens_i_qvec(qvec state) {
  for (PetscInt i = 0; i < state->n; i++) {
    PetscScalar v = 0.0;
    get_wf_element_ens_i_qvec_local(state, state->ens_i, i, &v);
    printf("%e + %ei\n", creal(v), cimag(v));
  }
}

/* qvec_utilities.c:1553-1589 (rows from kron.c:209-266): q <- M q for a
 * wavefunction (every trajectory of a spawned ensemble), rho <- M rho M^dag for
 * a density matrix (the superoperator conj(M) (x) M on the column-major vector),
 * with M = I (x) mat (x) I on the operator's subsystem and mat[r][c] its row r,
 * column c.  The reference keeps at most two entries per row of mat; any mat is
 * taken here.  The matrix is built on the host and applied by qvec_mat_mult. */
void apply_op_to_qvec_mat(qvec q, PetscScalar **mat, operator op) {
  const PetscInt N = q->total_levels, L = op->my_levels;
  const PetscInt na = N / (L * op->n_before);
  if (L < 1 || na < 1 || L * op->n_before * na != N) QUAC_FAIL(9, "apply_op_to_qvec_mat: the operator does not fit the state");
  const int dm = q->my_type == DENSITY_MATRIX;
  const PetscInt D = dm ? N * N : N, per = dm ? L * L : L;
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
  int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D * per));
  double complex *val = (double complex *)malloc(sizeof(double complex) * (size_t)(D * per));
  int64_t *mc = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N * L));   /* rows of M: columns */
  double complex *mv = (double complex *)malloc(sizeof(double complex) * (size_t)(N * L));
  PetscInt *mn = (PetscInt *)malloc(sizeof(PetscInt) * (size_t)N);
  if (!rp || !col || !val || !mc || !mv || !mn) {
    free(rp), free(col), free(val), free(mc), free(mv), free(mn);
    QUAC_FAIL(9, "out of host memory");
  }
  for (PetscInt i = 0; i < N; i++) {
    const PetscInt is = (i / na) % L;
    mn[i] = 0;
    for (PetscInt j = 0; j < L; j++)
      if (mat[is][j] != 0.0) {
        mc[L * i + mn[i]] = i + (j - is) * na;
        mv[L * i + mn[i]++] = mat[is][j];
      }
  }
  int64_t nnz = 0;
  for (PetscInt i = 0; i < D; i++) {
    rp[i] = nnz;
    if (!dm) {
      for (PetscInt k = 0; k < mn[i]; k++) col[nnz] = mc[L * i + k], val[nnz++] = mv[L * i + k];
    } else {
      const PetscInt i1 = i / N, i2 = i % N;
      for (PetscInt k1 = 0; k1 < mn[i1]; k1++)
        for (PetscInt k2 = 0; k2 < mn[i2]; k2++) {
          col[nnz] = N * mc[L * i1 + k1] + mc[L * i2 + k2];
          val[nnz++] = conj(mv[L * i1 + k1]) * mv[L * i2 + k2];
        }
    }
  }
  rp[D] = nnz;
  Mat A = NULL;
  int rc = 0;
  if (nnz == 0) { /* mat = 0: the state becomes 0 */
    rc = q->my_type == WF_ENSEMBLE && q->ens_spawned ? qh_vec_zero(q->ens_datas[0]) : qh_vec_zero(q->data);
  } else {
    rc = qh_mat_from_csr(D, rp, col, (const double *)val, &A);
  }
  free(rp), free(col), free(val), free(mc), free(mv), free(mn);
  if (rc) QUAC_FAIL(9, "apply_op_to_qvec_mat: %s", qh_last_error());
  if (A) {
    qvec_mat_mult(A, q);
    qh_mat_destroy(A);
  }
}
