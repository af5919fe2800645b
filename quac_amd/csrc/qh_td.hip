// qh_td.hip -- the time-dependent right-hand side as one fused kernel.
//
// The reference evaluates y = A x + sum_k s_k(t) A_k x with K + 1 MatMults and K
// axpys per RHS (qsystem.c:1288-1330; s_k = a_k f_k(t), times -i for a
// wavefunction, terms with |s_k| <= 1e-10 skipped), where A's pattern already
// holds every A_k's entries (seeded with a = 0, qsystem.c:645-653).  Here one
// kernel per RK stage reads A's CSR and a second, merged CSR of all the A_k
// entries (column, term index, value) and forms the row sum of both in FP64,
// with the RK4 stage epilogue fused (K_s and the next stage input, or the
// update).  The s_k of the four stages of a step are uploaded together once per
// step; the stage buffers live with the merged structure, which is built once
// per (A, A_1..A_K) and cached on A, so a step allocates nothing.
#include <cmath>
#include <cstring>
#include <vector>

#include "qh_common.h"
#include "qh_ks.h"

namespace {

const int TD_MAX = 150;  // time-dependent terms (the reference allocates 150)

// row sum of A x + sum_e s[k_e] v_e x[c_e] for one row, lanes of a G-group striding
template <typename T, int G>
__device__ inline double2 td_rowsum(int32_t row, int lane, const int32_t *__restrict__ rp,
                                    const int32_t *__restrict__ col, const T *__restrict__ val,
                                    const int32_t *__restrict__ trp, const int32_t *__restrict__ tcol,
                                    const int32_t *__restrict__ tk, const T *__restrict__ tval,
                                    const double2 *__restrict__ s, const T *__restrict__ x) {
  double re = 0.0, im = 0.0;
  for (int32_t k = rp[row] + lane; k < rp[row + 1]; k += G) {
    const T a = val[k], v = x[col[k]];
    re += (double)a.x * (double)v.x - (double)a.y * (double)v.y;
    im += (double)a.x * (double)v.y + (double)a.y * (double)v.x;
  }
  for (int32_t k = trp[row] + lane; k < trp[row + 1]; k += G) {
    const T a = tval[k], v = x[tcol[k]];
    const double2 c = s[tk[k]];
    // (c * a) * v with c = s_k(t)
    const double ar = c.x * (double)a.x - c.y * (double)a.y, ai = c.x * (double)a.y + c.y * (double)a.x;
    re += ar * (double)v.x - ai * (double)v.y;
    im += ar * (double)v.y + ai * (double)v.x;
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    re += __shfl_xor(re, off, G);
    im += __shfl_xor(im, off, G);
  }
  return make_double2(re, im);
}

template <typename T>
__device__ inline double2 ld_state(const T *x, const float2 *lo, int32_t i) {
  const T h = x[i];
  if (!lo) return make_double2((double)h.x, (double)h.y);
  const float2 l = lo[i];
  return make_double2((double)h.x + (double)l.x, (double)h.y + (double)l.y);
}

template <typename T>
__device__ inline void st_state(T *x, float2 *lo, int32_t i, double re, double im) {
  T o;
  o.x = re;
  o.y = im;
  x[i] = o;
  if (lo) lo[i] = make_float2((float)(re - (double)o.x), (float)(im - (double)o.y));
}

//   MODE_SPMV : kout = A xin + sum_k s_k A_k xin
//   MODE_STAGE: kout as above;  ynext = x + cn kout                (stages 1-3)
//   MODE_FINAL: x = x + h/6 (K1 + 2 K2 + 2 K3 + kout), in place    (stage 4)
// T = double2 (FP64) or float2 (FP32: x = hi + lo, lo != null)
template <typename T, int G, int MODE>
__global__ void __launch_bounds__(256) k_td_stage(int32_t nrows, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col, const T *__restrict__ val,
                                                  const int32_t *__restrict__ trp, const int32_t *__restrict__ tcol,
                                                  const int32_t *__restrict__ tk, const T *__restrict__ tval,
                                                  const double2 *__restrict__ s, const T *xin, T *kout, T *ynext,
                                                  double cn, T *x, float2 *lo, const T *k1, const T *k2,
                                                  const T *k3, double h) {
  const int64_t tid = (int64_t)qh_xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int32_t row = (int32_t)(tid / G);
  const int lane = (int)(tid % G);
  if (row >= nrows) return;  // G divides 256: a group never straddles the bound
  const double2 y = td_rowsum<T, G>(row, lane, rp, col, val, trp, tcol, tk, tval, s, xin);
  if (lane != 0) return;
  if (MODE == MODE_SPMV) {
    T o;
    o.x = y.x;
    o.y = y.y;
    kout[row] = o;
    return;
  }
  const double2 xo = ld_state(x, lo, row);
  if (MODE == MODE_STAGE) {
    T o;
    o.x = y.x;
    o.y = y.y;
    kout[row] = o;
    T yn;
    yn.x = xo.x + cn * y.x;
    yn.y = xo.y + cn * y.y;
    ynext[row] = yn;
    return;
  }
  const T a = k1[row], b = k2[row], c = k3[row];
  const double s6 = h * (1.0 / 6.0), s3 = h * (1.0 / 3.0);
  const double nr = xo.x + (s6 * (double)a.x + s3 * (double)b.x + s3 * (double)c.x + s6 * y.x);
  const double ni = xo.y + (s6 * (double)a.y + s3 * (double)b.y + s3 * (double)c.y + s6 * y.y);
  st_state(x, lo, row, nr, ni);
}

__global__ void k_narrow_td(const double2 *a, float2 *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

}  // namespace

// The merged structure of A_1..A_K and the stage buffers of one precision.
struct qh_td_cache {
  int ntd = 0;
  std::vector<const qh_mat *> ak;  // the matrices it was built from
  std::vector<uint64_t> akid;      // and their serial numbers
  int64_t rows = 0, tnnz = 0;
  int32_t max_row = 0;
  int32_t *trp = nullptr, *tcol = nullptr, *tk = nullptr;
  double2 *tval = nullptr;
  float2 *tval32 = nullptr;
  double2 *s = nullptr;       // 4 stages x TD_MAX coefficients (device)
  int prec = 0;               // precision of the stage buffers
  void *k[3] = {nullptr, nullptr, nullptr};
  void *y[2] = {nullptr, nullptr};
};

static void td_free_bufs(qh_td_cache *c) {
  for (auto &p : c->k) {
    qh_free_ptr(p);
    p = nullptr;
  }
  for (auto &p : c->y) {
    qh_free_ptr(p);
    p = nullptr;
  }
  c->prec = 0;
}

void qh_td_cache_free(qh_td_cache *c) {
  if (!c) return;
  (void)hipStreamSynchronize(qh_stream());
  qh_free_ptr(c->trp);
  qh_free_ptr(c->tcol);
  qh_free_ptr(c->tk);
  qh_free_ptr(c->tval);
  qh_free_ptr(c->tval32);
  qh_free_ptr(c->s);
  td_free_bufs(c);
  delete c;
}

// the merged CSR of the A_k (entries of each row in term order, then column
// order), built on the host from the device CSRs, cached on A.  Entries stored
// as exact zeros (every A_k carries the explicit zero diagonal of
// construct_matrix, qsystem.c:656-662) contribute s_k * 0 * x = 0 for a finite
// state and are left out: a tenth to a third of the entries for one-body drives.
static int td_cache_get(qh_mat *A, int ntd, qh_mat *const *Ak, qh_td_cache **out) {
  qh_td_cache *c = A->td_cache;
  bool ok = c && c->ntd == ntd;
  for (int k = 0; ok && k < ntd; k++) ok = c->ak[k] == Ak[k] && Ak[k] && c->akid[k] == Ak[k]->uid;
  if (ok) {
    *out = c;
    return QH_OK;
  }
  if (c) {
    qh_td_cache_free(c);
    A->td_cache = nullptr;
  }
  QH_REQUIRE(ntd >= 0 && ntd <= TD_MAX, QH_ERANGE, "time-dependent RHS: %d terms, at most %d", ntd, TD_MAX);
  int rc = qh_mat_max_row(A);
  if (rc) return rc;
  const int64_t n = A->rows;
  std::vector<std::vector<int64_t>> rps(ntd), cols(ntd);
  std::vector<std::vector<double>> vals(ntd);
  int64_t tot = 0;
  for (int k = 0; k < ntd; k++) {
    QH_REQUIRE(Ak[k] && Ak[k]->rows == n, QH_EINVAL, "time-dependent matrix %d does not match A", k);
    rps[k].resize(n + 1);
    cols[k].resize(Ak[k]->nnz);
    vals[k].resize(2 * Ak[k]->nnz);
    rc = qh_mat_export(Ak[k], rps[k].data(), cols[k].data(), vals[k].data());
    if (rc) return rc;
    tot += Ak[k]->nnz;
  }
  QH_REQUIRE(tot < ((int64_t)1 << 31), QH_ERANGE, "time-dependent matrices: %lld entries", (long long)tot);
  std::vector<int32_t> trp(n + 1), tcol(tot > 0 ? tot : 1), tk(tot > 0 ? tot : 1);
  std::vector<double2> tval(tot > 0 ? tot : 1);
  int64_t e = 0;
  int32_t mr = 0;
  for (int64_t r = 0; r < n; r++) {
    trp[r] = (int32_t)e;
    for (int k = 0; k < ntd; k++)
      for (int64_t j = rps[k][r]; j < rps[k][r + 1]; j++) {
        const double vr = vals[k][2 * j], vi = vals[k][2 * j + 1];
        if (vr == 0.0 && vi == 0.0) continue;
        tcol[e] = (int32_t)cols[k][j];
        tk[e] = k;
        tval[e] = make_double2(vr, vi);
        e++;
      }
    if ((int32_t)(e - trp[r]) > mr) mr = (int32_t)(e - trp[r]);
  }
  trp[n] = (int32_t)e;
  tot = e;
  tcol.resize(e > 0 ? e : 1);
  tk.resize(e > 0 ? e : 1);
  tval.resize(e > 0 ? e : 1);
  c = new qh_td_cache();
  c->ntd = ntd;
  c->ak.assign(Ak, Ak + ntd);
  c->akid.resize(ntd);
  for (int k = 0; k < ntd; k++) c->akid[k] = Ak[k]->uid;
  c->rows = n;
  c->tnnz = tot;
  c->max_row = mr;
  rc = qh_alloc_t(&c->trp, (size_t)n + 1);
  if (!rc) rc = qh_alloc_t(&c->tcol, tcol.size());
  if (!rc) rc = qh_alloc_t(&c->tk, tk.size());
  if (!rc) rc = qh_alloc_t(&c->tval, tval.size());
  if (!rc) rc = qh_alloc_t(&c->tval32, tval.size());
  if (!rc) rc = qh_alloc_t(&c->s, (size_t)4 * TD_MAX);
  if (rc) {
    qh_td_cache_free(c);
    return rc;
  }
  hipStream_t s = qh_stream();
  QH_HIP(hipMemcpyAsync(c->trp, trp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(c->tcol, tcol.data(), sizeof(int32_t) * tcol.size(), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(c->tk, tk.data(), sizeof(int32_t) * tk.size(), hipMemcpyHostToDevice, s));
  QH_HIP(hipMemcpyAsync(c->tval, tval.data(), sizeof(double2) * tval.size(), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_narrow_td, dim3(qh_blocks((int64_t)tval.size(), 256)), dim3(256), 0, s, c->tval,
                     c->tval32, (int64_t)tval.size());
  QH_HIP(hipGetLastError());
  QH_HIP(hipStreamSynchronize(s));
  A->td_cache = c;
  *out = c;
  return QH_OK;
}

// coefficients s_k with |s_k| <= 1e-10 set to zero (min_time_dep, qsystem.c:34,1320)
static void td_coefs(int ntd, const double *sri, double2 *dst) {
  for (int k = 0; k < ntd; k++) {
    const double re = sri[2 * k], im = sri[2 * k + 1];
    dst[k] = std::sqrt(re * re + im * im) <= 1e-10 ? make_double2(0.0, 0.0) : make_double2(re, im);
  }
}

template <int MODE>
static int td_launch(const qh_mat *A, const qh_td_cache *c, int prec, const double2 *s, const void *xin, void *kout,
                     void *ynext, double cn, void *x, float2 *lo, const void *k1, const void *k2, const void *k3,
                     double h) {
  const int64_t n = A->rows;
  const int longest = A->max_row > c->max_row ? A->max_row : c->max_row;
  const int G = longest > 24 ? 16 : 4;
  const dim3 grid(qh_blocks(n * G, 256));
  hipStream_t st = qh_stream();
#define QH_TDL(T, GG, V, TV)                                                                                          \
  hipLaunchKernelGGL((k_td_stage<T, GG, MODE>), grid, dim3(256), 0, st, (int32_t)n, A->rowptr, A->col, (const T *)V, \
                     c->trp, c->tcol, c->tk, (const T *)TV, s, (const T *)xin, (T *)kout, (T *)ynext, cn, (T *)x,   \
                     lo, (const T *)k1, (const T *)k2, (const T *)k3, h)
  if (prec == QH_FP64) {
    if (G == 16) QH_TDL(double2, 16, A->val, c->tval); else QH_TDL(double2, 4, A->val, c->tval);
  } else {
    if (G == 16) QH_TDL(float2, 16, A->val32, c->tval32); else QH_TDL(float2, 4, A->val32, c->tval32);
  }
#undef QH_TDL
  QH_HIP(hipGetLastError());
  return QH_OK;
}

extern "C" int qh_spmv_td(const qh_mat *A, int ntd, qh_mat *const *Ak, const double *s_re_im, const qh_vec *x,
                          qh_vec *y) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && x && y && (ntd == 0 || (Ak && s_re_im)), QH_EINVAL, "qh_spmv_td: null argument");
  QH_REQUIRE(x->n == A->rows && y->n == A->rows && x->prec == y->prec, QH_EINVAL,
             "qh_spmv_td: vectors of the matrix dimension and one precision required");
  QH_REQUIRE(x->prec == QH_FP64 || A->val32, QH_EINVAL, "qh_spmv_td: FP32 needs the complex64 values (qh_mat_make_fp32)");
  qh_td_cache *c = nullptr;
  rc = td_cache_get(const_cast<qh_mat *>(A), ntd, Ak, &c);
  if (rc) return rc;
  double2 hs[TD_MAX];
  td_coefs(ntd, s_re_im, hs);
  QH_HIP(hipMemcpyAsync(c->s, hs, sizeof(double2) * (ntd > 0 ? ntd : 1), hipMemcpyHostToDevice, qh_stream()));
  return td_launch<MODE_SPMV>(A, c, x->prec, c->s, x->d, y->d, nullptr, 0.0, nullptr, nullptr, nullptr, nullptr,
                              nullptr, 0.0);
}

extern "C" int qh_rk4_td_step(qh_mat *A, int ntd, qh_mat *const *Ak, const double *s4_re_im, qh_vec *x, double h) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && x && (ntd == 0 || (Ak && s4_re_im)), QH_EINVAL, "qh_rk4_td_step: null argument");
  QH_REQUIRE(x->n == A->rows, QH_EINVAL, "qh_rk4_td_step: state of the matrix dimension required");
  if (x->prec == QH_FP32 && !A->val32) {
    rc = qh_mat_make_fp32(A);
    if (rc) return rc;
  }
  qh_td_cache *c = nullptr;
  rc = td_cache_get(A, ntd, Ak, &c);
  if (rc) return rc;
  if (c->prec != x->prec) {
    td_free_bufs(c);
    const size_t es = x->prec == QH_FP64 ? 16 : 8;
    for (auto &p : c->k)
      if (!rc) rc = qh_alloc(&p, (size_t)A->rows * es);
    for (auto &p : c->y)
      if (!rc) rc = qh_alloc(&p, (size_t)A->rows * es);
    if (rc) {
      td_free_bufs(c);
      return rc;
    }
    c->prec = x->prec;
  }
  // the four stages' coefficients in one upload: stage j at s + j * TD_MAX
  double2 hs[4 * TD_MAX];
  for (int j = 0; j < 4; j++) td_coefs(ntd, s4_re_im + 2 * (size_t)j * ntd, hs + j * TD_MAX);
  QH_HIP(hipMemcpyAsync(c->s, hs, sizeof(double2) * 4 * TD_MAX, hipMemcpyHostToDevice, qh_stream()));
  float2 *lo = x->prec == QH_FP32 ? x->lo : nullptr;
  // K1 = f(t, x); Ya = x + h/2 K1; K2 = f(t + h/2, Ya); Yb = x + h/2 K2; K3 = f(t + h/2, Yb);
  // Ya = x + h K3; x += h/6 (K1 + 2 K2 + 2 K3 + f(t + h, Ya))
  const double cs[3] = {0.5 * h, 0.5 * h, h};
  const void *ins[3] = {x->d, c->y[0], c->y[1]};
  void *outs[3] = {c->y[0], c->y[1], c->y[0]};
  for (int j = 0; j < 3 && !rc; j++)
    rc = td_launch<MODE_STAGE>(A, c, x->prec, c->s + j * TD_MAX, ins[j], c->k[j], outs[j], cs[j], x->d, lo, nullptr,
                               nullptr, nullptr, 0.0);
  if (!rc)
    rc = td_launch<MODE_FINAL>(A, c, x->prec, c->s + 3 * TD_MAX, c->y[0], nullptr, nullptr, 0.0, x->d, lo, c->k[0],
                               c->k[1], c->k[2], h);
  return rc;
}
