/*
 * qo.c -- CPU oracle (test infrastructure only; see qo.h).
 *
 * Complex arithmetic is written out with the same operation order as the C99
 * complex operators the reference uses (a*b = (ar*br - ai*bi, ar*bi + ai*br)),
 * and this file is compiled with -ffp-contract=off so no FMA changes a bit.
 */
#include "qo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cx;

static inline cx cmk(double r, double i) { cx z; z.re = r; z.im = i; return z; }
static inline cx cmul(cx a, cx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cadd(cx a, cx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cx cconj(cx a) { return cmk(a.re, -a.im); }
static inline cx crmul(double s, cx a) { return cmk(s * a.re, s * a.im); }
static const cx C_I = {0.0, 1.0};

/* ------------------------------------------------------------------------- */
/* single operator row rule: kron.c:294-538 (_get_val_j_from_global_i)        */
/* ------------------------------------------------------------------------- */
static int single_op(int64_t tot, int64_t i, const qo_op *op, int tc, cx *val, int64_t *j)
{
  if (tc != QO_TENSOR_GG) {
    int64_t extra = (tc == QO_TENSOR_GI) ? tot : 1;
    int64_t L = op->levels;
    int64_t n_after = tot / (L * op->n_before) * extra;
    int64_t i_sub = i / n_after % L;
    int64_t tmp, k1, k2;
    *j = -1; *val = cmk(0.0, 0.0);
    switch (op->type) {
    case QO_LOWER:                                   /* kron.c:320-339 */
      if (i_sub < L - 1) {
        tmp = i - i_sub * n_after; k2 = tmp / (L * n_after); k1 = tmp % (L * n_after);
        *j = (i_sub + 1) * n_after + k1 + k2 * L * n_after;
        *val = cmk(sqrt((double)i_sub + 1.0), 0.0);
      }
      break;
    case QO_RAISE:                                   /* kron.c:340-360 */
      i_sub = i_sub - 1;
      if (i_sub >= 0) {
        tmp = i - (i_sub + 1) * n_after; k2 = tmp / (L * n_after); k1 = tmp % (L * n_after);
        *j = i_sub * n_after + k1 + k2 * L * n_after;
        *val = cmk(sqrt((double)i_sub + 1.0), 0.0);
      }
      break;
    case QO_SIGMA_X:                                 /* kron.c:361-393 */
    case QO_SIGMA_Y:                                 /* kron.c:394-431 */
      if (i_sub == 0 || i_sub == 1) {
        int64_t other = 1 - i_sub;
        tmp = i - i_sub * n_after; k2 = tmp / (L * n_after); k1 = tmp % (L * n_after);
        *j = other * n_after + k1 + k2 * L * n_after;
        if (op->type == QO_SIGMA_X) *val = cmk(1.0, 0.0);
        else *val = (i_sub == 0) ? cmk(-0.0, -1.0) : cmk(0.0, 1.0); /* -PETSC_i, PETSC_i */
      } else if (op->type == QO_SIGMA_Y) {
        return QO_ERR_PAULI;                         /* kron.c:424-429 exit(0) */
      }
      break;
    case QO_NUMBER:                                  /* kron.c:433-442 */
      if (i_sub != 0) { *j = i; *val = cmk((double)i_sub, 0.0); }
      break;
    case QO_IDENTITY:                                /* kron.c:443-447 */
      *j = i; *val = cmk(1.0, 0.0);
      break;
    case QO_SIGMA_Z:                                 /* kron.c:448-470 */
      if (i_sub == 0) { *j = i; *val = cmk(1.0, 0.0); }
      else if (i_sub == 1) { *j = i; *val = cmk(-1.0, 0.0); }
      else return QO_ERR_PAULI;
      break;
    case QO_OTHER:                                   /* kron.c:471-486 */
      if (i_sub >= 2) { *j = i; *val = cmk(1.0, 0.0); }
      break;
    default:
      return QO_ERR_VEC;                             /* kron.c:487-501 exit(0) */
    }
    if (tc == QO_TENSOR_GI) *val = cconj(*val);      /* kron.c:503-506 */
    return QO_OK;
  } else {
    /* U* x U: kron.c:507-537 */
    int64_t i1 = i / tot, i2 = i % tot, j1, j2;
    cx v1, v2;
    int rc = single_op(tot, i1, op, QO_TENSOR_IG, &v1, &j1);
    if (rc) return rc;
    rc = single_op(tot, i2, op, QO_TENSOR_IG, &v2, &j2);
    if (rc) return rc;
    if (j1 == -1 || j2 == -1) { *j = -1; *val = cmk(0.0, 0.0); }
    else { *j = tot * j1 + j2; *val = cmul(cconj(v1), v2); }
    return QO_OK;
  }
}

/* VEC pair |p1><p2|: kron.c:540-635 (_get_val_j_from_global_i_vec_vec) */
static int vec_vec(int64_t tot, int64_t i, const qo_op *o1, const qo_op *o2, int tc, cx *val, int64_t *j)
{
  if (o1->type != QO_VEC || o2->type != QO_VEC) return QO_ERR_VEC;
  if (o1->n_before != o2->n_before) return QO_ERR_VEC;
  if (tc != QO_TENSOR_GG) {
    int64_t extra = (tc == QO_TENSOR_GI) ? tot : 1;
    int64_t L = o1->levels;
    int64_t n_after = tot / (L * o1->n_before) * extra;
    int64_t i_sub = i / n_after % L;
    if (i_sub == o1->position) {
      int64_t tmp = i - i_sub * n_after, k2 = tmp / (L * n_after), k1 = tmp % (L * n_after);
      *j = o2->position * n_after + k1 + k2 * L * n_after;
      *val = cmk(1.0, 0.0);
    } else { *j = -1; *val = cmk(0.0, 0.0); }
    if (tc == QO_TENSOR_GI) *val = cconj(*val);
    return QO_OK;
  } else {
    int64_t i1 = i / tot, i2 = i % tot, j1, j2;
    cx v1, v2;
    vec_vec(tot, i1, o1, o2, QO_TENSOR_IG, &v1, &j1);
    vec_vec(tot, i2, o1, o2, QO_TENSOR_IG, &v2, &j2);
    if (j1 == -1 || j2 == -1) { *j = -1; *val = cmk(0.0, 0.0); }
    else { *j = tot * j1 + j2; *val = cmul(cconj(v1), v2); }
    return QO_OK;
  }
}

/* _get_val_j_ops, kron.c:161-207: apply ops[0], ops[1], ... to index *j. */
static int chain(int64_t tot, int nops, const qo_op *ops, int tc, cx *val_out, int64_t *j_inout)
{
  int64_t this_j = *j_inout, j2;
  cx val = cmk(1.0, 0.0), tmp;
  int rc;
  for (int k = 0; k < nops; k++) {
    if (ops[k].type == QO_VEC) {
      if (k + 1 >= nops || ops[k + 1].type != QO_VEC) return QO_ERR_VEC;
      if (this_j != -1) {
        rc = vec_vec(tot, this_j, &ops[k], &ops[k + 1], tc, &tmp, &j2);
        if (rc) return rc;
        this_j = j2; val = cmul(tmp, val);
      }
      k++;
    } else if (this_j != -1) {
      rc = single_op(tot, this_j, &ops[k], tc, &tmp, &j2);
      if (rc) return rc;
      this_j = j2; val = cmul(tmp, val);
    }
  }
  *val_out = val; *j_inout = this_j;
  return QO_OK;
}

int qo_chain(int64_t tot, int64_t i, int nops, const qo_op *ops, int tc,
             double *val_re, double *val_im, int64_t *j)
{
  cx v; int64_t jj = i;
  int rc = chain(tot, nops, ops, tc, &v, &jj);
  *val_re = v.re; *val_im = v.im; *j = jj;
  return rc;
}

/* ------------------------------------------------------------------------- */
/* PETSc-like AIJ builder: MatSetValue(ADD_VALUES) into sorted rows           */
/* ------------------------------------------------------------------------- */
typedef struct {
  int64_t n;
  int64_t *cap, *len, *start;   /* per-row capacity, length, start */
  int64_t *col;
  cx *val;
} aij;

/* MatSetValues_SeqAIJ: existing entry -> value added; new entry -> value set. */
static void aij_add(aij *A, int64_t r, int64_t c, cx v)
{
  int64_t s = A->start[r], n = A->len[r];
  int64_t lo = 0, hi = n;
  while (lo < hi) {                                  /* lower_bound(col) */
    int64_t mid = (lo + hi) / 2;
    if (A->col[s + mid] < c) lo = mid + 1; else hi = mid;
  }
  if (lo < n && A->col[s + lo] == c) { A->val[s + lo] = cadd(A->val[s + lo], v); return; }
  for (int64_t k = n; k > lo; k--) { A->col[s + k] = A->col[s + k - 1]; A->val[s + k] = A->val[s + k - 1]; }
  A->col[s + lo] = c; A->val[s + lo] = v; A->len[r] = n + 1;
}

typedef struct {
  int64_t tot, dim;
  int dm, mcwf, conj_a;
} sysinfo;

/* contributions per term, following the dispatch of kron.c:953-993 */
enum { K_HAM_DM, K_LIN_DM, K_HAM_ONLY, K_LIN_MCWF };

static int term_kind(const sysinfo *S, int64_t kind)
{
  int is_lin = (kind == QO_LIN || kind == QO_TD_LIN);
  if (S->dm) {
    if (S->mcwf) return is_lin ? K_LIN_MCWF : K_HAM_ONLY;
    return is_lin ? K_LIN_DM : K_HAM_DM;
  }
  if (is_lin) return -1;                              /* kron.c:985-990 exit */
  return K_HAM_ONLY;
}

/* dagger list for MCWF: kron.c:675-686 */
static qo_op *mcwf_ops(int nops, const qo_op *ops)
{
  qo_op *t = (qo_op *)malloc(sizeof(qo_op) * 2 * (size_t)nops);
  for (int i = 0; i < nops; i++) {
    qo_op d = ops[i];
    if (d.type == QO_LOWER) d.type = QO_RAISE;
    else if (d.type == QO_RAISE) d.type = QO_LOWER;
    t[nops - i - 1] = (ops[i].type == QO_VEC) ? ops[i] : d;
    t[nops + i] = ops[i];
  }
  return t;
}

/*
 * Apply one term to A (count_only: just count contributions per target row).
 * kron.c:638-657 (ham_only), :664-700 (lin_mcwf), :702-729 (ham), :995-1036 (lin).
 */
static int apply_term(const sysinfo *S, aij *A, int64_t *count, int tk, cx a, int nops, const qo_op *ops)
{
  int64_t tot = S->tot, n = S->dim;
  int rc = QO_OK;
  qo_op *mops = NULL;
  if (tk == K_LIN_MCWF) { mops = mcwf_ops(nops, ops); ops = mops; nops *= 2; }
  for (int64_t i = 0; i < n && rc == QO_OK; i++) {
    cx v_ig, v_gi, v_gg, add;
    int64_t j_ig = i, j_gi = i, j_gg = i;
    switch (tk) {
    case K_HAM_ONLY:
      rc = chain(tot, nops, ops, QO_TENSOR_IG, &v_ig, &j_ig);
      if (rc) break;
      if (j_ig != -1) {
        if (count) count[i]++;
        else { add = cmul(a, v_ig); aij_add(A, i, j_ig, add); }
      }
      break;
    case K_LIN_MCWF:
      rc = chain(tot, nops, ops, QO_TENSOR_IG, &v_ig, &j_ig);
      if (rc) break;
      if (j_ig != -1) {
        if (count) count[i]++;
        else {
          /* -PETSC_i*a/2*val_ig */
          cx t = cmul(cmk(-0.0, -1.0), a);
          t = cmk(t.re / 2, t.im / 2);
          add = cmul(t, v_ig);
          aij_add(A, i, j_ig, add);
        }
      }
      break;
    case K_HAM_DM:
      rc = chain(tot, nops, ops, QO_TENSOR_IG, &v_ig, &j_ig);
      if (rc) break;
      rc = chain(tot, nops, ops, QO_TENSOR_GI, &v_gi, &j_gi);
      if (rc) break;
      if (j_ig != -1) {
        if (count) count[i]++;
        else { add = cmul(cmul(cmk(-a.re, -a.im), C_I), v_ig); aij_add(A, i, j_ig, add); }
      }
      if (j_gi != -1) {
        if (count) count[j_gi]++;
        else {
          cx ac = S->conj_a ? cconj(a) : a;
          add = cmul(cmul(ac, C_I), cconj(v_gi));
          aij_add(A, j_gi, i, add);
        }
      }
      break;
    case K_LIN_DM:
      rc = chain(tot, nops, ops, QO_TENSOR_IG, &v_ig, &j_ig);
      if (rc) break;
      rc = chain(tot, nops, ops, QO_TENSOR_GI, &v_gi, &j_gi);
      if (rc) break;
      rc = chain(tot, nops, ops, QO_TENSOR_GG, &v_gg, &j_gg);
      if (rc) break;
      if (j_ig != -1) {
        if (count) count[j_ig]++;
        else { add = cmul(cmul(crmul(-0.5, a), cconj(v_ig)), v_ig); aij_add(A, j_ig, j_ig, add); }
      }
      if (j_gi != -1) {
        if (count) count[j_gi]++;
        else { add = cmul(crmul(-0.5, a), cconj(cmul(cconj(v_gi), v_gi))); aij_add(A, j_gi, j_gi, add); }
      }
      if (j_gg != -1) {
        if (count) count[i]++;
        else { add = cmul(a, v_gg); aij_add(A, i, j_gg, add); }
      }
      break;
    default:
      rc = QO_ERR_ARG;
    }
  }
  free(mops);
  return rc;
}

int qo_construct(int64_t tot, int dm, int mcwf, int conj_a, int nterms,
                 const int64_t *kind, const int64_t *nops, const int64_t *off,
                 const double *a, const int64_t *ops_flat, int td_index, qo_csr **out)
{
  sysinfo S;
  aij A;
  int rc = QO_OK;
  if (tot <= 0 || nterms <= 0) return QO_ERR_ARG;
  S.tot = tot; S.dm = dm; S.mcwf = mcwf; S.conj_a = conj_a;
  S.dim = (dm && !mcwf) ? tot * tot : tot;            /* qsystem.c:612-626 */
  const qo_op *ops = (const qo_op *)ops_flat;

  memset(&A, 0, sizeof(A));
  A.n = S.dim;
  int64_t *cnt = (int64_t *)calloc((size_t)S.dim, sizeof(int64_t));
  if (!cnt) return QO_ERR_MEM;
  /* order of insertion: time-indep terms, then time-dep terms with a=0
   * (qsystem.c:631-653), then an explicit 0 on the diagonal (:656-662). */
  for (int pass = 0; pass < 2 && rc == QO_OK; pass++) {
    aij *AA = pass ? &A : NULL;
    int64_t *cc = pass ? NULL : cnt;
    if (td_index >= 0) {
      int t = td_index;
      int tk = term_kind(&S, kind[t]);
      if (tk < 0) { rc = QO_ERR_ARG; break; }
      rc = apply_term(&S, AA, cc, tk, cmk(1.0, 0.0), (int)nops[t], ops + off[t]);
    } else {
      for (int t = 0; t < nterms && rc == QO_OK; t++) {
        if (kind[t] < 0) continue;
        int tk = term_kind(&S, kind[t]);
        if (tk < 0) { rc = QO_ERR_ARG; break; }
        rc = apply_term(&S, AA, cc, tk, cmk(a[2 * t], a[2 * t + 1]), (int)nops[t], ops + off[t]);
      }
      for (int t = 0; t < nterms && rc == QO_OK; t++) {
        if (kind[t] >= 0) continue;
        int tk = term_kind(&S, kind[t]);
        if (tk < 0) { rc = QO_ERR_ARG; break; }
        rc = apply_term(&S, AA, cc, tk, cmk(0.0, 0.0), (int)nops[t], ops + off[t]);
      }
    }
    if (rc) break;
    if (!pass) {
      int64_t total = 0;
      A.cap = cnt; A.start = (int64_t *)malloc(sizeof(int64_t) * (size_t)S.dim);
      A.len = (int64_t *)calloc((size_t)S.dim, sizeof(int64_t));
      for (int64_t r = 0; r < S.dim; r++) { cnt[r] += 1; A.start[r] = total; total += cnt[r]; }
      A.col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(total ? total : 1));
      A.val = (cx *)malloc(sizeof(cx) * (size_t)(total ? total : 1));
      if (!A.start || !A.len || !A.col || !A.val) { rc = QO_ERR_MEM; break; }
    } else {
      for (int64_t r = 0; r < S.dim; r++) aij_add(&A, r, r, cmk(0.0, 0.0));
    }
  }
  if (rc) {
    free(cnt); free(A.start); free(A.len); free(A.col); free(A.val);
    return rc;
  }
  /* compress into CSR */
  qo_csr *m = (qo_csr *)malloc(sizeof(qo_csr));
  m->n = S.dim;
  m->rowptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(S.dim + 1));
  int64_t nnz = 0;
  for (int64_t r = 0; r < S.dim; r++) { m->rowptr[r] = nnz; nnz += A.len[r]; }
  m->rowptr[S.dim] = nnz;
  m->nnz = nnz;
  m->col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  m->val = (double *)malloc(sizeof(double) * 2 * (size_t)(nnz ? nnz : 1));
  int scale = (td_index < 0) && (!dm || mcwf);        /* MatScale(A,-i), qsystem.c:670-673 */
  for (int64_t r = 0; r < S.dim; r++) {
    for (int64_t k = 0; k < A.len[r]; k++) {
      int64_t d = m->rowptr[r] + k;
      cx v = A.val[A.start[r] + k];
      if (scale) v = cmul(cmk(-0.0, -1.0), v);        /* zscal with alpha = -i */
      m->col[d] = A.col[A.start[r] + k];
      m->val[2 * d] = v.re; m->val[2 * d + 1] = v.im;
    }
  }
  free(cnt); free(A.start); free(A.len); free(A.col); free(A.val);
  *out = m;
  return QO_OK;
}

int64_t qo_csr_rows(const qo_csr *m) { return m->n; }
int64_t qo_csr_nnz(const qo_csr *m) { return m->nnz; }
void qo_csr_export(const qo_csr *m, int64_t *rowptr, int64_t *col, double *val)
{
  if (rowptr) memcpy(rowptr, m->rowptr, sizeof(int64_t) * (size_t)(m->n + 1));
  if (col) memcpy(col, m->col, sizeof(int64_t) * (size_t)m->nnz);
  if (val) memcpy(val, m->val, sizeof(double) * 2 * (size_t)m->nnz);
}
void qo_csr_free(qo_csr *m)
{
  if (!m) return;
  free(m->rowptr); free(m->col); free(m->val); free(m);
}

/* ------------------------------------------------------------------------- */
/* MatMult and RK integrators (TS wrapper, qsystem.c:895-1098)                */
/* ------------------------------------------------------------------------- */
void qo_spmv(const qo_csr *m, const double *x, double *y)
{
  int64_t n = m->n;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    double sr = 0.0, si = 0.0;
    for (int64_t k = m->rowptr[r]; k < m->rowptr[r + 1]; k++) {
      double ar = m->val[2 * k], ai = m->val[2 * k + 1];
      double xr = x[2 * m->col[k]], xi = x[2 * m->col[k] + 1];
      sr += ar * xr - ai * xi;
      si += ar * xi + ai * xr;
    }
    y[2 * r] = sr; y[2 * r + 1] = si;
  }
}

/* Y = X + c*K (complex vectors, real c) */
static void axpy_to(int64_t n2, double *Y, const double *X, double c, const double *K)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n2; i++) Y[i] = X[i] + c * K[i];
}

int qo_rk4(const qo_csr *m, double *x, double dt, int64_t nsteps, int nthreads)
{
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  int64_t n2 = 2 * m->n;
  double *k1 = (double *)malloc(sizeof(double) * (size_t)n2);
  double *k2 = (double *)malloc(sizeof(double) * (size_t)n2);
  double *k3 = (double *)malloc(sizeof(double) * (size_t)n2);
  double *k4 = (double *)malloc(sizeof(double) * (size_t)n2);
  double *y = (double *)malloc(sizeof(double) * (size_t)n2);
  if (!k1 || !k2 || !k3 || !k4 || !y) return QO_ERR_MEM;
  const double b1 = 1.0 / 6.0, b2 = 1.0 / 3.0;
  for (int64_t s = 0; s < nsteps; s++) {
    qo_spmv(m, x, k1);
    axpy_to(n2, y, x, 0.5 * dt, k1);
    qo_spmv(m, y, k2);
    axpy_to(n2, y, x, 0.5 * dt, k2);
    qo_spmv(m, y, k3);
    axpy_to(n2, y, x, dt, k3);
    qo_spmv(m, y, k4);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n2; i++)
      x[i] = x[i] + dt * (b1 * k1[i] + b2 * k2[i] + b2 * k3[i] + b1 * k4[i]);
  }
  free(k1); free(k2); free(k3); free(k4); free(y);
  return QO_OK;
}

/* y = A x + sum_k s_k(t) A_k x, s_k = a_k f_k(t) (x -i for WF): qsystem.c:1300-1330 */
static void td_rhs(const qo_csr *m, int ntd, qo_csr *const *td, const double *td_a, int wf_scale,
                   qo_tdfunc f, double t, const double *x, double *y, double *w)
{
  int64_t n2 = 2 * m->n;
  qo_spmv(m, x, y);
  for (int k = 0; k < ntd; k++) {
    double fr, fi;
    f(t, k, &fr, &fi);
    cx s = cmul(cmk(td_a[2 * k], td_a[2 * k + 1]), cmk(fr, fi));
    if (wf_scale) s = cmul(cmk(-0.0, -1.0), s);
    if (sqrt(s.re * s.re + s.im * s.im) > 1e-10) {   /* min_time_dep, qsystem.c:34,1320 */
      qo_spmv(td[k], x, w);
      for (int64_t i = 0; i < n2; i += 2) {
        y[i] += s.re * w[i] - s.im * w[i + 1];
        y[i + 1] += s.re * w[i + 1] + s.im * w[i];
      }
    }
  }
}

int qo_rk4_td(const qo_csr *m, int ntd, qo_csr *const *td, const double *td_a, int wf_scale,
              qo_tdfunc f, double *x, double t0, double dt, int64_t nsteps)
{
  int64_t n2 = 2 * m->n;
  double *k[4], *y = (double *)malloc(sizeof(double) * (size_t)n2), *w = (double *)malloc(sizeof(double) * (size_t)n2);
  for (int i = 0; i < 4; i++) k[i] = (double *)malloc(sizeof(double) * (size_t)n2);
  double t = t0;
  for (int64_t s = 0; s < nsteps; s++) {
    td_rhs(m, ntd, td, td_a, wf_scale, f, t, x, k[0], w);
    axpy_to(n2, y, x, 0.5 * dt, k[0]);
    td_rhs(m, ntd, td, td_a, wf_scale, f, t + 0.5 * dt, y, k[1], w);
    axpy_to(n2, y, x, 0.5 * dt, k[1]);
    td_rhs(m, ntd, td, td_a, wf_scale, f, t + 0.5 * dt, y, k[2], w);
    axpy_to(n2, y, x, dt, k[2]);
    td_rhs(m, ntd, td, td_a, wf_scale, f, t + dt, y, k[3], w);
    for (int64_t i = 0; i < n2; i++)
      x[i] = x[i] + dt * ((1.0 / 6.0) * k[0][i] + (1.0 / 3.0) * k[1][i] + (1.0 / 3.0) * k[2][i] + (1.0 / 6.0) * k[3][i]);
    t = t0 + (double)(s + 1) * dt;
  }
  for (int i = 0; i < 4; i++) free(k[i]);
  free(y); free(w);
  return QO_OK;
}

/*
 * One Bogacki-Shampine 3(2) trial step (PETSc TSRK3BS tableau, qsystem.c:1004;
 * Bogacki & Shampine 1989): k1 = A x (taken as given when have_k1: FSAL),
 * k2 = A(x + h/2 k1), k3 = A(x + 3h/4 k2), xn = x + h(2/9 k1 + 1/3 k2 + 4/9 k3),
 * k4 = A xn, and the embedded error e = h((2/9-7/24) k1 + (1/3-1/4) k2 +
 * (4/9-1/3) k3 - 1/8 k4) measured as the weighted RMS norm
 * sqrt(mean_i |e_i|^2 / (atol + rtol max(|x_i|, |xn_i|))^2).
 */
int qo_rk3bs_trial(const qo_csr *m, const double *x, double h, int have_k1, double *k1, double *k2,
                   double *k3, double *k4, double *xn, double atol, double rtol, double *err)
{
  int64_t n2 = 2 * m->n;
  double *y = malloc(sizeof(double) * n2);
  if (!y) return QO_ERR_MEM;
  if (!have_k1) qo_spmv(m, x, k1);
  axpy_to(n2, y, x, 0.5 * h, k1);
  qo_spmv(m, y, k2);
  axpy_to(n2, y, x, 0.75 * h, k2);
  qo_spmv(m, y, k3);
  for (int64_t i = 0; i < n2; i++)
    xn[i] = x[i] + h * ((2.0 / 9.0) * k1[i] + (1.0 / 3.0) * k2[i] + (4.0 / 9.0) * k3[i]);
  qo_spmv(m, xn, k4);
  double e2 = 0.0;
  for (int64_t i = 0; i < n2; i += 2) {
    double er = h * ((2.0 / 9.0 - 7.0 / 24.0) * k1[i] + (1.0 / 3.0 - 1.0 / 4.0) * k2[i] +
                     (4.0 / 9.0 - 1.0 / 3.0) * k3[i] - (1.0 / 8.0) * k4[i]);
    double ei = h * ((2.0 / 9.0 - 7.0 / 24.0) * k1[i + 1] + (1.0 / 3.0 - 1.0 / 4.0) * k2[i + 1] +
                     (4.0 / 9.0 - 1.0 / 3.0) * k3[i + 1] - (1.0 / 8.0) * k4[i + 1]);
    double ax = sqrt(x[i] * x[i] + x[i + 1] * x[i + 1]);
    double an = sqrt(xn[i] * xn[i] + xn[i + 1] * xn[i + 1]);
    double sc = atol + rtol * (ax > an ? ax : an);
    e2 += (er * er + ei * ei) / (sc * sc);
  }
  *err = sqrt(e2 / (double)m->n);
  free(y);
  return QO_OK;
}

/*
 * Adaptive Bogacki-Shampine 3(2) with a basic controller around qo_rk3bs_trial:
 * h_new = h * clamp(0.9 * err^(-1/3), 0.1, 10) limited to [hmin, hmax], and the
 * last step clipped to land on tmax (TS_EXACTFINALTIME_MATCHSTEP, qsystem.c:1001).
 */
int qo_rk3bs(const qo_csr *m, double *x, double t0, double tmax, double dt0,
             int64_t steps_max, double atol, double rtol, double hmin, double hmax,
             int64_t *steps_taken, double *t_end)
{
  int64_t n2 = 2 * m->n;
  double *k1 = malloc(sizeof(double) * n2), *k2 = malloc(sizeof(double) * n2);
  double *k3 = malloc(sizeof(double) * n2), *k4 = malloc(sizeof(double) * n2);
  double *xn = malloc(sizeof(double) * n2);
  double t = t0, h = dt0;
  int64_t steps = 0;
  int have_k1 = 0;
  while (t < tmax - 1e-12 && steps < steps_max) {
    if (t + h > tmax) h = tmax - t;
    double err = 0.0;
    int rc = qo_rk3bs_trial(m, x, h, have_k1, k1, k2, k3, k4, xn, atol, rtol, &err);
    if (rc) return rc;
    double fac = 0.9 * pow(err > 1e-300 ? err : 1e-300, -1.0 / 3.0);
    if (fac < 0.1) fac = 0.1;
    if (fac > 10.0) fac = 10.0;
    double hnew = h * fac;
    if (hnew < hmin) hnew = hmin;
    if (hnew > hmax) hnew = hmax;
    if (err <= 1.0 || h <= hmin * (1 + 1e-12)) {
      memcpy(x, xn, sizeof(double) * n2);
      memcpy(k1, k4, sizeof(double) * n2);
      have_k1 = 1;
      t += h;
      steps++;
    }
    h = hnew;
  }
  if (steps_taken) *steps_taken = steps;
  if (t_end) *t_end = t;
  free(k1); free(k2); free(k3); free(k4); free(xn);
  return QO_OK;
}

/* ------------------------------------------------------------------------- */
/* observables: qvec_utilities.c:1197-1281                                    */
/* ------------------------------------------------------------------------- */
int qo_expect_dm(int64_t tot, const double *x, int nops, const qo_op *ops, double *re, double *im)
{
  cx acc = cmk(0.0, 0.0);
  for (int64_t i = 0; i < tot; i++) {
    int64_t k = i;
    cx v;
    int rc = chain(tot, nops, ops, QO_TENSOR_IG, &v, &k);
    if (rc) return rc;
    if (k < 0) continue;
    int64_t loc = tot * i + k;                        /* rho[k, i], :1266 */
    acc = cadd(acc, cmul(v, cmk(x[2 * loc], x[2 * loc + 1])));
  }
  *re = acc.re; *im = acc.im;
  return QO_OK;
}

int qo_expect_wf(int64_t tot, const double *x, int nops, const qo_op *ops, double *re, double *im)
{
  cx acc = cmk(0.0, 0.0);
  for (int64_t i = 0; i < tot; i++) {
    int64_t k = i;
    cx v;
    int rc = chain(tot, nops, ops, QO_TENSOR_IG, &v, &k);
    if (rc) return rc;
    if (k < 0) continue;
    /* conj(psi_i) * O_ik * psi_k */
    cx t = cmul(v, cmk(x[2 * k], x[2 * k + 1]));
    acc = cadd(acc, cmul(cmk(x[2 * i], -x[2 * i + 1]), t));
  }
  *re = acc.re; *im = acc.im;
  return QO_OK;
}

/* ------------------------------------------------------------------------- */
/* gates                                                                      */
/* ------------------------------------------------------------------------- */
/* 2x2 unitary of a single-qubit gate in QuaC's definition */
static int gate_u2(int gate, double theta, double phi, double lambda, cx u[2][2])
{
  double s2 = pow(2.0, -0.5);
  switch (gate) {
  case QO_HADAMARD:                                 /* quantum_circuits.c:2149-2236 */
    u[0][0] = cmk(s2, 0); u[0][1] = cmk(s2, 0); u[1][0] = cmk(s2, 0); u[1][1] = cmk(-s2, 0); break;
  case QO_SIGMAX:
    u[0][0] = cmk(0, 0); u[0][1] = cmk(1, 0); u[1][0] = cmk(1, 0); u[1][1] = cmk(0, 0); break;
  case QO_SIGMAY:
    u[0][0] = cmk(0, 0); u[0][1] = cmk(0, -1); u[1][0] = cmk(0, 1); u[1][1] = cmk(0, 0); break;
  case QO_SIGMAZ:
    u[0][0] = cmk(1, 0); u[0][1] = cmk(0, 0); u[1][0] = cmk(0, 0); u[1][1] = cmk(-1, 0); break;
  case QO_EYE:
    u[0][0] = cmk(1, 0); u[0][1] = cmk(0, 0); u[1][0] = cmk(0, 0); u[1][1] = cmk(1, 0); break;
  case QO_RX:                                       /* :2650-2739 */
    u[0][0] = cmk(cos(theta / 2), 0); u[0][1] = cmk(0, -sin(theta / 2));
    u[1][0] = cmk(0, -sin(theta / 2)); u[1][1] = cmk(cos(theta / 2), 0); break;
  case QO_RY:
    u[0][0] = cmk(cos(theta / 2), 0); u[0][1] = cmk(-sin(theta / 2), 0);
    u[1][0] = cmk(sin(theta / 2), 0); u[1][1] = cmk(cos(theta / 2), 0); break;
  case QO_RZ:                                       /* :2484-2558, diag(1, e^{i theta}) */
    u[0][0] = cmk(1, 0); u[0][1] = cmk(0, 0); u[1][0] = cmk(0, 0); u[1][1] = cmk(cos(theta), sin(theta)); break;
  case QO_U1: case QO_U2: case QO_U3: {             /* U3_gate_func_wf, :2055-2110 */
    double c = cos(theta / 2), s = sin(theta / 2);
    u[0][0] = cmk(c, 0);
    u[0][1] = cmk(-cos(lambda) * s, -sin(lambda) * s);
    u[1][0] = cmk(cos(phi) * s, sin(phi) * s);
    u[1][1] = cmk(cos(lambda + phi) * c, sin(lambda + phi) * c);
    break; }
  default:
    return QO_ERR_ARG;
  }
  return QO_OK;
}

/* 4x4 unitary for 2-qubit gates, basis |q0 q1> with q0 the first qubit arg (control) */
static int gate_u4(int gate, cx u[4][4])
{
  memset(u, 0, sizeof(cx) * 16);
  switch (gate) {
  case QO_CNOT:
    u[0][0] = u[1][1] = cmk(1, 0); u[2][3] = u[3][2] = cmk(1, 0); break;
  case QO_CZ:
    u[0][0] = u[1][1] = u[2][2] = cmk(1, 0); u[3][3] = cmk(-1, 0); break;
  case QO_CXZ: /* I2 (+) sig_x sig_z, quantum_circuits.c:1076-1080 */
    u[0][0] = u[1][1] = cmk(1, 0); u[2][3] = cmk(-1, 0); u[3][2] = cmk(1, 0); break;
  default:
    return QO_ERR_ARG;
  }
  return QO_OK;
}

/*
 * Apply an operator acting on one or two subsystems (strides s0[,s1], levels 2)
 * of an index space of size n, to vector x with element stride `xs` ... generic
 * dense application: for each basis index i, y_i = sum_j U[sub(i)][sub(j)] x_j.
 */
static void apply_local(int64_t n, int nq, const int64_t *stride, const cx *U, int conj, double *x)
{
  int64_t dimU = (nq == 1) ? 2 : 4;
  double *y = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int64_t i = 0; i < n; i++) {
    int64_t si = 0, base = i;
    for (int q = 0; q < nq; q++) {
      int64_t b = (i / stride[q]) % 2;
      si = si * 2 + b;
      base -= b * stride[q];
    }
    cx acc = cmk(0.0, 0.0);
    for (int64_t sj = 0; sj < dimU; sj++) {
      int64_t j = base;
      for (int q = 0; q < nq; q++) {
        int64_t b = (sj >> (nq - 1 - q)) & 1;
        j += b * stride[q];
      }
      cx u = U[si * dimU + sj];
      if (conj) u = cconj(u);
      acc = cadd(acc, cmul(u, cmk(x[2 * j], x[2 * j + 1])));
    }
    y[2 * i] = acc.re; y[2 * i + 1] = acc.im;
  }
  memcpy(x, y, sizeof(double) * 2 * (size_t)n);
  free(y);
}

int qo_gate_superop(int nsub, const int64_t *levels, const int64_t *nbefore, int dm,
                    int gate, const int64_t *qubits, double theta, double phi, double lambda,
                    double *x)
{
  int64_t tot = 1;
  for (int s = 0; s < nsub; s++) tot *= levels[s];
  int nq = gate < 0 ? 2 : 1;
  cx U[16];
  int rc = (nq == 1) ? gate_u2(gate, theta, phi, lambda, (cx(*)[2])U) : gate_u4(gate, (cx(*)[4])U);
  if (rc) return rc;
  int64_t stride[2];
  for (int q = 0; q < nq; q++) {
    int64_t sub = qubits[q];
    if (sub < 0 || sub >= nsub || levels[sub] != 2) return QO_ERR_ARG;
    stride[q] = tot / (levels[sub] * nbefore[sub]);  /* n_after: _get_n_after_1qbit_sys, :2982-3004 */
  }
  if (!dm) { apply_local(tot, nq, stride, U, 0, x); return QO_OK; }
  /* DM: rows (fast index) get U, columns (slow index) get U*  -> U rho U^dagger */
  int64_t n = tot * tot;
  apply_local(n, nq, stride, U, 0, x);
  int64_t cs[2] = {stride[0] * tot, nq == 2 ? stride[1] * tot : 0};
  apply_local(n, nq, cs, U, 1, x);
  return QO_OK;
}

int qo_gate_inplace(int nsub, const int64_t *dims, int dm, int gate, const int64_t *qubits,
                    double theta, double phi, double lambda, int rz_conj_bug, double *x)
{
  int64_t tot = 1;
  for (int s = 0; s < nsub; s++) tot *= dims[s];
  int nq = gate < 0 ? 2 : 1;
  cx U[16];
  int rc = (nq == 1) ? gate_u2(gate, theta, phi, lambda, (cx(*)[2])U) : gate_u4(gate, (cx(*)[4])U);
  if (rc) return rc;
  int64_t stride[2];
  for (int q = 0; q < nq; q++) {                     /* n_bef = prod dims[0..q-1], :1824-1842 */
    int64_t sb = 1;
    if (qubits[q] < 0 || qubits[q] >= nsub) return QO_ERR_ARG;
    for (int k = 0; k < qubits[q]; k++) sb *= dims[k];
    stride[q] = sb;
  }
  int64_t n = dm ? tot * tot : tot;
  apply_local(n, nq, stride, U, 0, x);
  if (dm) {
    int64_t cs[2] = {stride[0] * tot, nq == 2 ? stride[1] * tot : 0};
    int conj = !(gate == QO_RZ && rz_conj_bug);
    apply_local(n, nq, cs, U, conj, x);
  }
  return QO_OK;
}

/* ------------------------------------------------------------------------- */
/* steady state: solver.c:33-220 semantics, restarted GMRES (FP64)            */
/* ------------------------------------------------------------------------- */
static void stab_matvec(const qo_csr *m, int64_t tot, const double *dinv, const double *v, double *y, double *tmp)
{
  /* right preconditioning: y = A' D^{-1} v */
  const double *z = v;
  if (dinv) {
    for (int64_t i = 0; i < m->n; i++) {
      tmp[2 * i] = dinv[2 * i] * v[2 * i] - dinv[2 * i + 1] * v[2 * i + 1];
      tmp[2 * i + 1] = dinv[2 * i] * v[2 * i + 1] + dinv[2 * i + 1] * v[2 * i];
    }
    z = tmp;
  }
  qo_spmv(m, z, y);
  double tr = 0.0, ti = 0.0;
  for (int64_t i = 0; i < tot; i++) { tr += z[2 * i * (tot + 1)]; ti += z[2 * i * (tot + 1) + 1]; }
  y[0] += tr; y[1] += ti;                              /* row 0 += 1 at cols i(N+1), :60-72 */
}

static double nrm2(int64_t n, const double *v)
{
  double s = 0.0;
  for (int64_t i = 0; i < 2 * n; i++) s += v[i] * v[i];
  return sqrt(s);
}

int qo_gmres(const qo_csr *m, int64_t tot, double *x, int restart, double rtol_abs,
             int64_t maxit, int jacobi, int64_t *iters, double *res)
{
  int64_t n = m->n, it = 0;
  int R = restart;
  double *V = (double *)malloc(sizeof(double) * 2 * (size_t)n * (R + 1));
  double *w = (double *)malloc(sizeof(double) * 2 * n), *tmp = (double *)malloc(sizeof(double) * 2 * n);
  double *r = (double *)malloc(sizeof(double) * 2 * n);
  cx *H = (cx *)calloc((size_t)(R + 1) * R, sizeof(cx));
  cx *cs = (cx *)malloc(sizeof(cx) * R), *sn = (cx *)malloc(sizeof(cx) * R), *g = (cx *)malloc(sizeof(cx) * (R + 1));
  double *dinv = NULL;
  if (jacobi) {
    dinv = (double *)malloc(sizeof(double) * 2 * n);
    for (int64_t i = 0; i < n; i++) {
      cx d = cmk(0, 0);
      for (int64_t k = m->rowptr[i]; k < m->rowptr[i + 1]; k++)
        if (m->col[k] == i) d = cmk(m->val[2 * k], m->val[2 * k + 1]);
      if (i == 0) d.re += 1.0;                         /* stabilisation touches (0,0) */
      double dd = d.re * d.re + d.im * d.im;
      if (dd == 0.0) { dinv[2 * i] = 1.0; dinv[2 * i + 1] = 0.0; }
      else { dinv[2 * i] = d.re / dd; dinv[2 * i + 1] = -d.im / dd; }
    }
  }
  double rn = 0.0;
  /* the preconditioned unknown u satisfies x = D^{-1} u; start from u = D x */
  double *u = (double *)malloc(sizeof(double) * 2 * n);
  if (dinv) {
    for (int64_t i = 0; i < n; i++) {            /* u = D x = x / dinv */
      double dr = dinv[2 * i], di = dinv[2 * i + 1], dd = dr * dr + di * di;
      double xr = x[2 * i], xi = x[2 * i + 1];
      u[2 * i] = (xr * dr + xi * di) / dd; u[2 * i + 1] = (xi * dr - xr * di) / dd;
    }
  } else memcpy(u, x, sizeof(double) * 2 * n);
  for (;;) {
    stab_matvec(m, tot, dinv, u, r, tmp);
    for (int64_t i = 0; i < 2 * n; i++) r[i] = -r[i];
    r[0] += 1.0;                                       /* b = e0 */
    rn = nrm2(n, r);
    if (rn <= rtol_abs || it >= maxit) break;
    for (int64_t i = 0; i < 2 * n; i++) V[i] = r[i] / rn;
    for (int k = 0; k <= R; k++) g[k] = cmk(0, 0);
    g[0] = cmk(rn, 0);
    int k;
    for (k = 0; k < R && it < maxit; k++, it++) {
      double *vk = V + 2 * n * k, *vn = V + 2 * n * (k + 1);
      stab_matvec(m, tot, dinv, vk, w, tmp);
      for (int pass = 0; pass < 2; pass++) {          /* CGS2 */
        for (int jj = 0; jj <= k; jj++) {
          double *vj = V + 2 * n * jj;
          double hr = 0, hi = 0;
          for (int64_t i = 0; i < n; i++) {           /* h = <vj, w> = sum conj(vj) w */
            hr += vj[2 * i] * w[2 * i] + vj[2 * i + 1] * w[2 * i + 1];
            hi += vj[2 * i] * w[2 * i + 1] - vj[2 * i + 1] * w[2 * i];
          }
          H[jj * R + k] = cadd(H[jj * R + k], cmk(hr, hi));
          for (int64_t i = 0; i < n; i++) {
            w[2 * i] -= hr * vj[2 * i] - hi * vj[2 * i + 1];
            w[2 * i + 1] -= hr * vj[2 * i + 1] + hi * vj[2 * i];
          }
        }
      }
      double hn = nrm2(n, w);
      H[(k + 1) * R + k] = cmk(hn, 0);
      for (int64_t i = 0; i < 2 * n; i++) vn[i] = hn > 0 ? w[i] / hn : 0.0;
      for (int jj = 0; jj < k; jj++) {                /* apply previous Givens */
        cx a = H[jj * R + k], b = H[(jj + 1) * R + k];
        H[jj * R + k] = cadd(cmul(cconj(cs[jj]), a), cmul(cconj(sn[jj]), b));
        H[(jj + 1) * R + k] = cadd(cmul(cmk(-sn[jj].re, -sn[jj].im), a), cmul(cs[jj], b));
      }
      cx a = H[k * R + k], b = H[(k + 1) * R + k];
      double na = sqrt(a.re * a.re + a.im * a.im), nb = sqrt(b.re * b.re + b.im * b.im);
      double den = sqrt(na * na + nb * nb);
      /* rotation [c* s*; -s c] with c = a/den, s = b/den maps (a,b) -> (den,0) */
      if (den == 0) { cs[k] = cmk(1, 0); sn[k] = cmk(0, 0); }
      else { cs[k] = cmk(a.re / den, a.im / den); sn[k] = cmk(b.re / den, b.im / den); }
      cx gk = g[k];
      g[k] = cmul(cconj(cs[k]), gk);
      g[k + 1] = cmul(cmk(-sn[k].re, -sn[k].im), gk);
      H[k * R + k] = cadd(cmul(cconj(cs[k]), a), cmul(cconj(sn[k]), b));
      H[(k + 1) * R + k] = cmk(0, 0);
      double gn = sqrt(g[k + 1].re * g[k + 1].re + g[k + 1].im * g[k + 1].im);
      if (gn <= rtol_abs) { k++; it++; break; }
    }
    /* back substitution */
    cx *yv = (cx *)malloc(sizeof(cx) * (size_t)(k > 0 ? k : 1));
    for (int i = k - 1; i >= 0; i--) {
      cx s = g[i];
      for (int jj = i + 1; jj < k; jj++) { cx t = cmul(H[i * R + jj], yv[jj]); s = cmk(s.re - t.re, s.im - t.im); }
      cx d = H[i * R + i];
      double dd = d.re * d.re + d.im * d.im;
      yv[i] = cmk((s.re * d.re + s.im * d.im) / dd, (s.im * d.re - s.re * d.im) / dd);
    }
    for (int jj = 0; jj < k; jj++) {
      double *vj = V + 2 * n * jj;
      for (int64_t i = 0; i < n; i++) {
        u[2 * i] += yv[jj].re * vj[2 * i] - yv[jj].im * vj[2 * i + 1];
        u[2 * i + 1] += yv[jj].re * vj[2 * i + 1] + yv[jj].im * vj[2 * i];
      }
    }
    free(yv);
    memset(H, 0, sizeof(cx) * (size_t)(R + 1) * R);
  }
  if (dinv) {
    for (int64_t i = 0; i < n; i++) {
      x[2 * i] = dinv[2 * i] * u[2 * i] - dinv[2 * i + 1] * u[2 * i + 1];
      x[2 * i + 1] = dinv[2 * i] * u[2 * i + 1] + dinv[2 * i + 1] * u[2 * i];
    }
  } else memcpy(x, u, sizeof(double) * 2 * n);
  if (iters) *iters = it;
  if (res) *res = rn;
  free(V); free(w); free(tmp); free(r); free(H); free(cs); free(sn); free(g); free(dinv); free(u);
  return QO_OK;
}
