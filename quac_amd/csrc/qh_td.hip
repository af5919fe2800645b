// qh_td.hip -- the time-dependent right-hand side.
//
// The reference evaluates y = A x + sum_k s_k(t) A_k x with K + 1 MatMults and K
// axpys per RHS (qsystem.c:1288-1330; s_k = a_k f_k(t), times -i for a
// wavefunction, terms with |s_k| <= 1e-10 skipped), where A's pattern already
// holds every A_k's entries (seeded with a = 0, qsystem.c:645-653).  Here the
// entries of all the A_k are merged into one CSR (built once per (A, A_1..A_K),
// cached on A), and an RK stage is either
//   - A's own stage kernel on its autotuned sliced layout, summing the merged
//     entries of each row in the same kernel where the variant allows (the
//     one-row layouts share the 64-row slices; a row-pair lane's two rows are
//     adjacent in one of them), or followed by one pass over
//     the merged entries that adds sum_k s_k A_k y to the stage's K, its next
//     stage input or the state (k_td_ell, k_td_add), or
//   - one fused kernel over A's CSR and the merged CSR (k_td_stage), where A
//     has no sliced layout (small generators, CSR variants).
// The merged entries carry a column and either a 2-byte index into a dictionary
// of the distinct (term, value) pairs -- whose coefficients s_k(t) v are formed
// on the host per stage, a few per term for the usual drives -- or, past the
// dictionary's size limit, the term index and the value.  The coefficients of
// the four stages of a step are uploaded together once per step; the stage
// buffers live with the merged structure, so a step allocates nothing.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "qh_common.h"
#include "qh_ks.h"

namespace {

const int TD_MAX = 150;  // time-dependent terms (the reference allocates 150)

// the merged A_k entries of one row, lanes of a G-group striding: with the
// dictionary (DICT) entry e adds cf[tidx_e] x[c_e], cf = the stage's s_k v table;
// without it s[k_e] v_e x[c_e]
template <typename T, int G, bool DICT>
__device__ inline void td_part(int32_t row, int lane, const int32_t *__restrict__ trp,
                               const int32_t *__restrict__ tcol, const int32_t *__restrict__ tk,
                               const T *__restrict__ tval, const uint16_t *__restrict__ tidx,
                               const double2 *__restrict__ s, const T *__restrict__ x, double &re, double &im) {
  for (int32_t k = trp[row] + lane; k < trp[row + 1]; k += G) {
    const T v = x[tcol[k]];
    double cr, ci;
    if (DICT) {
      const double2 c = s[tidx[k]];
      cr = c.x;
      ci = c.y;
    } else {
      const T a = tval[k];
      const double2 c = s[tk[k]];
      cr = c.x * (double)a.x - c.y * (double)a.y;
      ci = c.x * (double)a.y + c.y * (double)a.x;
    }
    re += cr * (double)v.x - ci * (double)v.y;
    im += cr * (double)v.y + ci * (double)v.x;
  }
}

template <int G>
__device__ inline void group_sum(double &re, double &im) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    re += __shfl_xor(re, off, G);
    im += __shfl_xor(im, off, G);
  }
}

// row sum of A x + the merged A_k entries for one row
template <typename T, int G, bool DICT>
__device__ inline double2 td_rowsum(int32_t row, int lane, const int32_t *__restrict__ rp,
                                    const int32_t *__restrict__ col, const T *__restrict__ val,
                                    const int32_t *__restrict__ trp, const int32_t *__restrict__ tcol,
                                    const int32_t *__restrict__ tk, const T *__restrict__ tval,
                                    const uint16_t *__restrict__ tidx, const double2 *__restrict__ s,
                                    const T *__restrict__ x) {
  double re = 0.0, im = 0.0;
  for (int32_t k = rp[row] + lane; k < rp[row + 1]; k += G) {
    const T a = val[k], v = x[col[k]];
    re += (double)a.x * (double)v.x - (double)a.y * (double)v.y;
    im += (double)a.x * (double)v.y + (double)a.y * (double)v.x;
  }
  td_part<T, G, DICT>(row, lane, trp, tcol, tk, tval, tidx, s, x, re, im);
  group_sum<G>(re, im);
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
template <typename T, int G, int MODE, bool DICT>
__global__ void __launch_bounds__(256) k_td_stage(int32_t nrows, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col, const T *__restrict__ val,
                                                  const int32_t *__restrict__ trp, const int32_t *__restrict__ tcol,
                                                  const int32_t *__restrict__ tk, const T *__restrict__ tval,
                                                  const uint16_t *__restrict__ tidx,
                                                  const double2 *__restrict__ s, const T *xin, T *kout, T *ynext,
                                                  double cn, T *x, float2 *lo, const T *k1, const T *k2,
                                                  const T *k3, double h) {
  const int64_t tid = (int64_t)qh_xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int32_t row = (int32_t)(tid / G);
  const int lane = (int)(tid % G);
  if (row >= nrows) return;  // G divides 256: a group never straddles the bound
  const double2 y = td_rowsum<T, G, DICT>(row, lane, rp, col, val, trp, tcol, tk, tval, tidx, s, xin);
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

// the A_k pass after A's own stage kernel has run on its sliced layout:
//   MODE_SPMV : kout += d                      with d = sum_k s_k A_k xin
//   MODE_STAGE: kout += d;  ynext += cn d      (stages 1-3)
//   MODE_FINAL: x += h/6 d, in place           (stage 4; FP32: x = hi + lo)
template <typename T, int G, int MODE, bool DICT>
__global__ void __launch_bounds__(256) k_td_add(int32_t nrows, const int32_t *__restrict__ trp,
                                                const int32_t *__restrict__ tcol, const int32_t *__restrict__ tk,
                                                const T *__restrict__ tval, const uint16_t *__restrict__ tidx,
                                                const double2 *__restrict__ s, const T *__restrict__ xin,
                                                T *kout, T *ynext, double cn, T *x, float2 *lo, double h) {
  const int64_t tid = (int64_t)qh_xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int32_t row = (int32_t)(tid / G);
  const int lane = (int)(tid % G);
  if (row >= nrows) return;  // G divides 256: a group never straddles the bound
  double re = 0.0, im = 0.0;
  td_part<T, G, DICT>(row, lane, trp, tcol, tk, tval, tidx, s, xin, re, im);
  group_sum<G>(re, im);
  if (lane != 0) return;
  if (MODE == MODE_SPMV || MODE == MODE_STAGE) {
    T k = kout[row];
    const double kr = (double)k.x + re, ki = (double)k.y + im;
    k.x = kr;
    k.y = ki;
    kout[row] = k;
    if (MODE == MODE_STAGE) {
      T yn = ynext[row];
      yn.x = (double)yn.x + cn * re;
      yn.y = (double)yn.y + cn * im;
      ynext[row] = yn;
    }
    return;
  }
  const double2 xo = ld_state(x, lo, row);
  const double s6 = h * (1.0 / 6.0);
  st_state(x, lo, row, xo.x + s6 * re, xo.y + s6 * im);
}

// The A_k pass over the sliced form of the dictionary entries: slices of 64
// rows, slice s holding w_s entries per row (the longest of its rows; shorter
// rows padded with column = the row itself and the dictionary's zero entry),
// entry i of row 64 s + l at soff_s + 64 i + l, so the one thread per row reads
// its columns and indices in fully coalesced wavefront loads.  Epilogues as
// k_td_add.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_td_ell(int32_t nrows, const int64_t *__restrict__ soff,
                                                const int32_t *__restrict__ sw, const int32_t *__restrict__ ecol,
                                                const uint16_t *__restrict__ eidx, const double2 *__restrict__ cf,
                                                const T *__restrict__ xin, T *kout, T *ynext, double cn, T *x,
                                                float2 *lo, double h) {
  const int32_t row = (int32_t)((int64_t)qh_xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
  if (row >= nrows) return;
  const int32_t sl = row >> 6;
  const int32_t w = sw[sl];
  const int64_t base = soff[sl] + (row & 63);
  double re = 0.0, im = 0.0;
#pragma unroll 4
  for (int32_t i = 0; i < w; i++) {
    const int64_t e = base + 64 * (int64_t)i;
    const T v = xin[ecol[e]];
    const double2 c = cf[eidx[e]];
    re += c.x * (double)v.x - c.y * (double)v.y;
    im += c.x * (double)v.y + c.y * (double)v.x;
  }
  if (MODE == MODE_SPMV || MODE == MODE_STAGE) {
    T k = kout[row];
    k.x = (double)k.x + re;
    k.y = (double)k.y + im;
    kout[row] = k;
    if (MODE == MODE_STAGE) {
      T yn = ynext[row];
      yn.x = (double)yn.x + cn * re;
      yn.y = (double)yn.y + cn * im;
      ynext[row] = yn;
    }
    return;
  }
  const double2 xo = ld_state(x, lo, row);
  const double s6 = h * (1.0 / 6.0);
  st_state(x, lo, row, xo.x + s6 * re, xo.y + s6 * im);
}

// nsteps classic RK4 steps of a generator of at most 800 rows with the
// time-dependent entries (dictionary form) in one workgroup: the state and the
// stage derivatives in LDS (FP64; an FP32 state is widened from hi + lo and
// split again at the end), stage s of step j taking the coefficient table
// cf + (4 j + s) st.  Each stage input is formed on the fly from the previous
// stage (x + c k), as in the one-workgroup RK4 of a time-independent generator.
template <bool F32>
__global__ void __launch_bounds__(1024) k_rk4_td_small(int32_t n, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ col,
                                                       const double2 *__restrict__ val,
                                                       const int32_t *__restrict__ trp,
                                                       const int32_t *__restrict__ tcol,
                                                       const uint16_t *__restrict__ tidx,
                                                       const double2 *__restrict__ cf, int32_t st, void *xg,
                                                       float2 *xlo, double h, int64_t nsteps) {
  extern __shared__ double2 sm[];
  double2 *xa = sm, *xb = sm + n, *k1 = sm + 2 * n, *k2 = sm + 3 * n, *k3 = sm + 4 * n;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int r = tid; r < n; r += nt) {
    if (F32) {
      const float2 hi = ((const float2 *)xg)[r], lo = xlo[r];
      xa[r] = make_double2((double)hi.x + (double)lo.x, (double)hi.y + (double)lo.y);
    } else {
      xa[r] = ((const double2 *)xg)[r];
    }
  }
  __syncthreads();
  auto rowdot = [&](int r, const double2 *xs, const double2 *kin, double c, const double2 *tab) {
    double are = 0.0, aim = 0.0;
    for (int32_t k = rp[r]; k < rp[r + 1]; k++) {
      const int32_t cc = col[k];
      double2 v = xs[cc];
      if (kin) {
        v.x += c * kin[cc].x;
        v.y += c * kin[cc].y;
      }
      const double2 a = val[k];
      are = fma(a.x, v.x, fma(-a.y, v.y, are));
      aim = fma(a.x, v.y, fma(a.y, v.x, aim));
    }
    for (int32_t k = trp[r]; k < trp[r + 1]; k++) {
      const int32_t cc = tcol[k];
      double2 v = xs[cc];
      if (kin) {
        v.x += c * kin[cc].x;
        v.y += c * kin[cc].y;
      }
      const double2 a = tab[tidx[k]];
      are = fma(a.x, v.x, fma(-a.y, v.y, are));
      aim = fma(a.x, v.y, fma(a.y, v.x, aim));
    }
    return make_double2(are, aim);
  };
  const double s6 = h * (1.0 / 6.0), s3 = h * (1.0 / 3.0);
  for (int64_t j = 0; j < nsteps; j++) {
    const double2 *tb = cf + 4 * j * (int64_t)st;
    for (int r = tid; r < n; r += nt) k1[r] = rowdot(r, xa, nullptr, 0.0, tb);
    __syncthreads();
    for (int r = tid; r < n; r += nt) k2[r] = rowdot(r, xa, k1, 0.5 * h, tb + st);
    __syncthreads();
    for (int r = tid; r < n; r += nt) k3[r] = rowdot(r, xa, k2, 0.5 * h, tb + 2 * st);
    __syncthreads();
    for (int r = tid; r < n; r += nt) {
      const double2 k4 = rowdot(r, xa, k3, h, tb + 3 * st), a = k1[r], b = k2[r], c = k3[r], xo = xa[r];
      xb[r] = make_double2(xo.x + s6 * a.x + s3 * b.x + s3 * c.x + s6 * k4.x,
                           xo.y + s6 * a.y + s3 * b.y + s3 * c.y + s6 * k4.y);
    }
    __syncthreads();
    double2 *t = xa;
    xa = xb;
    xb = t;
  }
  for (int r = tid; r < n; r += nt) {
    const double2 v = xa[r];
    if (F32) {
      const float2 hi = make_float2((float)v.x, (float)v.y);
      ((float2 *)xg)[r] = hi;
      xlo[r] = make_float2((float)(v.x - (double)hi.x), (float)(v.y - (double)hi.y));
    } else {
      ((double2 *)xg)[r] = v;
    }
  }
}

__global__ void k_narrow_td(const double2 *a, float2 *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

}  // namespace

// switches of qh_td_config: the dictionary's size limit (0: no dictionary), A's
// part on its sliced layout, the lanes per row of the A_k pass (0: from the
// merged rows' average length)
static int g_dict_max = 4096;
static const int TD_DICT_CAP = 8192;
// generators of at most this many rows take a run of RK4 steps in one workgroup
// (five state-sized FP64 arrays in LDS: 64 KB at 800 rows)
#define QH_TD_SMALL_ROWS 800
static int g_td_sliced = 1;
static int g_td_lanes = 0;

// The merged structure of A_1..A_K and the stage buffers of one precision.
struct qh_td_cache {
  int ntd = 0;
  std::vector<const qh_mat *> ak;  // the matrices it was built from
  std::vector<uint64_t> akid;      // and their serial numbers
  int64_t rows = 0, tnnz = 0;
  int32_t max_row = 0;
  int dict_max = 0;                 // the limit it was built under
  int32_t *trp = nullptr, *tcol = nullptr, *tk = nullptr;
  double2 *tval = nullptr;
  float2 *tval32 = nullptr;
  uint16_t *tidx = nullptr;         // dictionary form: index of the (term, value) pair
  int64_t *soff = nullptr;          // its sliced form (slices of 64 rows): slice offsets,
  int32_t *sw = nullptr;            // entries per row of each slice,
  int32_t *ecol = nullptr;          // columns
  uint16_t *eidx = nullptr;         // and dictionary indices (the zero entry: index dv.size() - 1)
  std::vector<int> dk;              // the dictionary's terms
  std::vector<double2> dv;          // and values (host)
  double2 *s = nullptr;             // 4 stages x (TD_MAX coefficients, or the dictionary's s_k v)
  double2 *steps_cf = nullptr;      // the stage tables of a run of steps (one-workgroup path)
  size_t steps_cap = 0;             // its capacity (tables)
  int prec = 0;                     // precision of the stage buffers
  void *k[3] = {nullptr, nullptr, nullptr};
  void *y[2] = {nullptr, nullptr};
  bool dict() const { return tidx != nullptr; }
  int stride() const { return dict() ? (int)dv.size() : TD_MAX; }  // coefficients per stage
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
  qh_free_ptr(c->tidx);
  qh_free_ptr(c->soff);
  qh_free_ptr(c->sw);
  qh_free_ptr(c->ecol);
  qh_free_ptr(c->eidx);
  qh_free_ptr(c->s);
  qh_free_ptr(c->steps_cf);
  td_free_bufs(c);
  delete c;
}

// the merged CSR of the A_k (entries of each row in term order, then column
// order), built on the host from the device CSRs, cached on A.  Entries stored
// as exact zeros (every A_k carries the explicit zero diagonal of
// construct_matrix, qsystem.c:656-662) contribute s_k * 0 * x = 0 for a finite
// state and are left out: a tenth to a third of the entries for one-body drives.
// The distinct (term, value) pairs form the dictionary when there are at most
// g_dict_max of them.
struct td_key {
  int k;
  uint64_t re, im;
  bool operator==(const td_key &o) const { return k == o.k && re == o.re && im == o.im; }
};
struct td_key_hash {
  size_t operator()(const td_key &a) const {
    uint64_t h = a.re * 0x9E3779B97F4A7C15ull;
    h ^= (a.im + 0x632BE59BD9B4E019ull) + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)a.k * 0xC2B2AE3D27D4EB4Full;
    return (size_t)(h ^ (h >> 31));
  }
};

static int td_cache_get(qh_mat *A, int ntd, qh_mat *const *Ak, qh_td_cache **out) {
  qh_td_cache *c = A->td_cache;
  bool ok = c && c->ntd == ntd && c->dict_max == g_dict_max;
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
  // the dictionary of distinct (term, value) pairs (bit patterns: -0.0 and 0.0
  // are two values, both exact)
  std::vector<uint16_t> tidx;
  std::vector<int> dk;
  std::vector<double2> dv;
  if (g_dict_max > 0 && tot > 0) {
    std::unordered_map<td_key, int, td_key_hash> idx;
    tidx.resize(tot);
    td_key last{-1, 0, 0};
    int last_i = -1;
    bool over = false;
    for (int64_t q = 0; q < tot && !over; q++) {
      td_key key{tk[q], 0, 0};
      std::memcpy(&key.re, &tval[q].x, 8);
      std::memcpy(&key.im, &tval[q].y, 8);
      if (!(key == last)) {
        auto it = idx.find(key);
        if (it == idx.end()) {
          if ((int)dk.size() >= g_dict_max) {
            over = true;
            break;
          }
          it = idx.emplace(key, (int)dk.size()).first;
          dk.push_back(key.k);
          dv.push_back(tval[q]);
        }
        last = key;
        last_i = it->second;
      }
      tidx[q] = (uint16_t)last_i;
    }
    if (over) {
      tidx.clear();
      dk.clear();
      dv.clear();
    } else {
      // the zero entry, for the padding of the sliced form
      dk.push_back(-1);
      dv.push_back(make_double2(0.0, 0.0));
    }
  }
  // the sliced form of the dictionary entries (skipped where padding would more
  // than double them)
  const int64_t nsl = (n + 63) / 64;
  std::vector<int64_t> soff;
  std::vector<int32_t> sw, ecol;
  std::vector<uint16_t> eidx;
  if (!dk.empty()) {
    soff.resize(nsl + 1);
    sw.resize(nsl > 0 ? nsl : 1);
    int64_t te = 0;
    for (int64_t q = 0; q < nsl; q++) {
      int32_t w = 0;
      for (int64_t r = 64 * q; r < std::min<int64_t>(n, 64 * q + 64); r++) w = std::max(w, trp[r + 1] - trp[r]);
      sw[q] = w;
      soff[q] = te;
      te += 64 * (int64_t)w;
    }
    soff[nsl] = te;
    if (te <= 2 * tot + 64 * nsl) {
      const uint16_t zero = (uint16_t)(dk.size() - 1);
      ecol.resize(te > 0 ? te : 1);
      eidx.resize(te > 0 ? te : 1);
      for (int64_t q = 0; q < nsl; q++)
        for (int64_t l = 0; l < 64; l++) {
          const int64_t r = 64 * q + l;
          const int32_t len = r < n ? trp[r + 1] - trp[r] : 0;
          for (int32_t i = 0; i < sw[q]; i++) {
            const int64_t e = soff[q] + 64 * (int64_t)i + l;
            ecol[e] = i < len ? tcol[trp[r] + i] : (int32_t)(r < n ? r : 0);
            eidx[e] = i < len ? tidx[trp[r] + i] : zero;
          }
        }
    } else {
      soff.clear();
    }
  }
  c = new qh_td_cache();
  c->ntd = ntd;
  c->ak.assign(Ak, Ak + ntd);
  c->akid.resize(ntd);
  for (int k = 0; k < ntd; k++) c->akid[k] = Ak[k]->uid;
  c->rows = n;
  c->tnnz = tot;
  c->max_row = mr;
  c->dict_max = g_dict_max;
  const size_t ne = tot > 0 ? (size_t)tot : 1;
  rc = qh_alloc_t(&c->trp, (size_t)n + 1);
  if (!rc) rc = qh_alloc_t(&c->tcol, ne);
  const bool dict = !dk.empty();
  if (dict) {
    c->dk = dk;
    c->dv = dv;
    if (!rc) rc = qh_alloc_t(&c->tidx, ne);
    if (!rc) rc = qh_alloc_t(&c->s, (size_t)4 * dv.size());
    if (!soff.empty()) {
      if (!rc) rc = qh_alloc_t(&c->soff, soff.size());
      if (!rc) rc = qh_alloc_t(&c->sw, sw.size());
      if (!rc) rc = qh_alloc_t(&c->ecol, ecol.size());
      if (!rc) rc = qh_alloc_t(&c->eidx, eidx.size());
    }
  } else {
    if (!rc) rc = qh_alloc_t(&c->tk, ne);
    if (!rc) rc = qh_alloc_t(&c->tval, ne);
    if (!rc) rc = qh_alloc_t(&c->tval32, ne);
    if (!rc) rc = qh_alloc_t(&c->s, (size_t)4 * TD_MAX);
  }
  if (rc) {
    qh_td_cache_free(c);
    return rc;
  }
  hipStream_t s = qh_stream();
  QH_HIP(hipMemcpyAsync(c->trp, trp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
  if (c->soff) {
    QH_HIP(hipMemcpyAsync(c->soff, soff.data(), sizeof(int64_t) * soff.size(), hipMemcpyHostToDevice, s));
    QH_HIP(hipMemcpyAsync(c->sw, sw.data(), sizeof(int32_t) * sw.size(), hipMemcpyHostToDevice, s));
    QH_HIP(hipMemcpyAsync(c->ecol, ecol.data(), sizeof(int32_t) * ecol.size(), hipMemcpyHostToDevice, s));
    QH_HIP(hipMemcpyAsync(c->eidx, eidx.data(), sizeof(uint16_t) * eidx.size(), hipMemcpyHostToDevice, s));
  }
  if (tot > 0) {
    QH_HIP(hipMemcpyAsync(c->tcol, tcol.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, s));
    if (dict) {
      QH_HIP(hipMemcpyAsync(c->tidx, tidx.data(), sizeof(uint16_t) * tot, hipMemcpyHostToDevice, s));
    } else {
      QH_HIP(hipMemcpyAsync(c->tk, tk.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, s));
      QH_HIP(hipMemcpyAsync(c->tval, tval.data(), sizeof(double2) * tot, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_narrow_td, dim3(qh_blocks(tot, 256)), dim3(256), 0, s, c->tval, c->tval32, tot);
      QH_HIP(hipGetLastError());
    }
  }
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

// the coefficient table of one stage: s_k per term, or s_k v per dictionary entry
static void td_table(const qh_td_cache *c, int ntd, const double *sri, double2 *dst) {
  double2 sk[TD_MAX];
  td_coefs(ntd, sri, sk);
  if (!c->dict()) {
    for (int k = 0; k < ntd; k++) dst[k] = sk[k];
    return;
  }
  for (size_t d = 0; d < c->dv.size(); d++) {
    if (c->dk[d] < 0) {  // the padding's zero entry
      dst[d] = make_double2(0.0, 0.0);
      continue;
    }
    const double2 a = sk[c->dk[d]], v = c->dv[d];
    dst[d] = make_double2(a.x * v.x - a.y * v.y, a.x * v.y + a.y * v.x);
  }
}

// the fused kernel: A's CSR and the merged A_k in one pass
template <int MODE>
static int td_launch(const qh_mat *A, const qh_td_cache *c, int prec, const double2 *s, const void *xin, void *kout,
                     void *ynext, double cn, void *x, float2 *lo, const void *k1, const void *k2, const void *k3,
                     double h) {
  const int64_t n = A->rows;
  const int longest = A->max_row > c->max_row ? A->max_row : c->max_row;
  const int G = longest > 24 ? 16 : 4;
  const dim3 grid(qh_blocks(n * G, 256));
  hipStream_t st = qh_stream();
#define QH_TDL(T, GG, V, TV, D)                                                                                     \
  hipLaunchKernelGGL((k_td_stage<T, GG, MODE, D>), grid, dim3(256), 0, st, (int32_t)n, A->rowptr, A->col,          \
                     (const T *)V, c->trp, c->tcol, c->tk, (const T *)TV, c->tidx, s, (const T *)xin, (T *)kout,   \
                     (T *)ynext, cn, (T *)x, lo, (const T *)k1, (const T *)k2, (const T *)k3, h)
#define QH_TDD(T, GG, V, TV) \
  if (c->dict()) QH_TDL(T, GG, V, TV, true); else QH_TDL(T, GG, V, TV, false)
  if (prec == QH_FP64) {
    if (G == 16) { QH_TDD(double2, 16, A->val, c->tval); } else { QH_TDD(double2, 4, A->val, c->tval); }
  } else {
    if (G == 16) { QH_TDD(float2, 16, A->val32, c->tval32); } else { QH_TDD(float2, 4, A->val32, c->tval32); }
  }
#undef QH_TDD
#undef QH_TDL
  QH_HIP(hipGetLastError());
  return QH_OK;
}

// the A_k pass after A's sliced stage kernel
template <int MODE>
static int td_add_launch(const qh_td_cache *c, int prec, const double2 *s, const void *xin, void *kout, void *ynext,
                         double cn, void *x, float2 *lo, double h) {
  const int64_t n = c->rows;
  hipStream_t st = qh_stream();
  if (c->soff && g_td_lanes <= 1) {  // the sliced form, one thread per row (a pass of its own)
    const dim3 grid(qh_blocks(n, 256));
    if (prec == QH_FP64)
      hipLaunchKernelGGL((k_td_ell<double2, MODE>), grid, dim3(256), 0, st, (int32_t)n, c->soff, c->sw, c->ecol,
                         c->eidx, s, (const double2 *)xin, (double2 *)kout, (double2 *)ynext, cn, (double2 *)x, lo, h);
    else
      hipLaunchKernelGGL((k_td_ell<float2, MODE>), grid, dim3(256), 0, st, (int32_t)n, c->soff, c->sw, c->ecol,
                         c->eidx, s, (const float2 *)xin, (float2 *)kout, (float2 *)ynext, cn, (float2 *)x, lo, h);
    QH_HIP(hipGetLastError());
    return QH_OK;
  }
  int G = g_td_lanes > 1 ? g_td_lanes : 0;
  if (G == 0) G = (double)c->tnnz <= 48.0 * (double)(n > 0 ? n : 1) ? 4 : 8;
  const dim3 grid(qh_blocks(n * G, 256));
#define QH_TDA(T, GG, TV, D)                                                                                        \
  hipLaunchKernelGGL((k_td_add<T, GG, MODE, D>), grid, dim3(256), 0, st, (int32_t)n, c->trp, c->tcol, c->tk,        \
                     (const T *)TV, c->tidx, s, (const T *)xin, (T *)kout, (T *)ynext, cn, (T *)x, lo, h)
#define QH_TDG(T, TV, D)                 \
  switch (G) {                           \
    case 4: QH_TDA(T, 4, TV, D); break;  \
    case 8: QH_TDA(T, 8, TV, D); break;  \
    default: QH_TDA(T, 16, TV, D); break; \
  }
  if (prec == QH_FP64) {
    if (c->dict()) { QH_TDG(double2, c->tval, true) } else { QH_TDG(double2, c->tval, false) }
  } else {
    if (c->dict()) { QH_TDG(float2, c->tval32, true) } else { QH_TDG(float2, c->tval32, false) }
  }
#undef QH_TDG
#undef QH_TDA
  QH_HIP(hipGetLastError());
  return QH_OK;
}

// the sliced form of the merged A_k for A's stage kernel to sum in the same pass
// (lanes 0: wherever A's variant has that form), or null
static const qh_td_ell *td_fusable(const qh_td_cache *c, const double2 *cf, qh_td_ell *buf) {
  if (!c->soff || g_td_lanes != 0) return nullptr;
  *buf = qh_td_ell{c->soff, c->sw, c->ecol, c->eidx, cf};
  return buf;
}

extern "C" int qh_td_config(int dict_max, int sliced, int lanes) {
  QH_REQUIRE(dict_max >= 0 && dict_max <= TD_DICT_CAP, QH_EINVAL, "qh_td_config: dictionary limit %d outside 0..%d",
             dict_max, TD_DICT_CAP);
  QH_REQUIRE(sliced == 0 || sliced == 1, QH_EINVAL, "qh_td_config: sliced must be 0 or 1");
  QH_REQUIRE(lanes == 0 || lanes == 1 || lanes == 4 || lanes == 8 || lanes == 16, QH_EINVAL,
             "qh_td_config: lanes per row must be 0, 1, 4, 8 or 16");
  g_dict_max = dict_max;
  g_td_sliced = sliced;
  g_td_lanes = lanes;
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
  QH_REQUIRE(x->d != y->d, QH_EINVAL, "qh_spmv_td: the input and output vectors must differ");
  qh_mat *Am = const_cast<qh_mat *>(A);
  qh_td_cache *c = nullptr;
  rc = td_cache_get(Am, ntd, Ak, &c);
  if (rc) return rc;
  int G = 0;
  if (g_td_sliced) {
    rc = qh_td_sliced_prepare(Am, x->prec, &G);
    if (rc) return rc;
  }
  std::vector<double2> hs((size_t)c->stride() + 1);
  td_table(c, ntd, s_re_im, hs.data());
  QH_HIP(hipMemcpyAsync(c->s, hs.data(), sizeof(double2) * c->stride(), hipMemcpyHostToDevice, qh_stream()));
  if (G) {
    qh_td_ell tb;
    int fused = 0;
    rc = qh_td_sliced_stage(G, A, x->prec, MODE_SPMV, x->d, y->d, nullptr, 0.0, nullptr, nullptr, nullptr, nullptr,
                            nullptr, 0.0, nullptr, td_fusable(c, c->s, &tb), &fused);
    if (rc || fused) return rc;
    return td_add_launch<MODE_SPMV>(c, x->prec, c->s, x->d, y->d, nullptr, 0.0, nullptr, nullptr, 0.0);
  }
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
  int G = 0;
  if (g_td_sliced) {
    rc = qh_td_sliced_prepare(A, x->prec, &G);
    if (rc) return rc;
  }
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
  // the four stages' coefficient tables in one upload: stage j at s + j * stride
  const int st = c->stride();
  std::vector<double2> hs((size_t)4 * st);
  for (int j = 0; j < 4; j++) td_table(c, ntd, s4_re_im + 2 * (size_t)j * ntd, hs.data() + (size_t)j * st);
  QH_HIP(hipMemcpyAsync(c->s, hs.data(), sizeof(double2) * 4 * st, hipMemcpyHostToDevice, qh_stream()));
  float2 *lo = x->prec == QH_FP32 ? x->lo : nullptr;
  // K1 = f(t, x); Ya = x + h/2 K1; K2 = f(t + h/2, Ya); Yb = x + h/2 K2; K3 = f(t + h/2, Yb);
  // Ya = x + h K3; x += h/6 (K1 + 2 K2 + 2 K3 + f(t + h, Ya))
  const double cs[3] = {0.5 * h, 0.5 * h, h};
  const void *ins[3] = {x->d, c->y[0], c->y[1]};
  void *outs[3] = {c->y[0], c->y[1], c->y[0]};
  if (G) {
    // A's part on its sliced layout with the A_k entries summed in the same
    // kernel where its variant allows, else followed by the A_k pass that adds
    // into the stage's outputs
    qh_td_ell tb;
    int fused = 0;
    for (int j = 0; j < 3 && !rc; j++) {
      const double2 *cf = c->s + (size_t)j * st;
      rc = qh_td_sliced_stage(G, A, x->prec, MODE_STAGE, ins[j], c->k[j], outs[j], cs[j], x->d, nullptr, nullptr,
                              nullptr, nullptr, h, lo, td_fusable(c, cf, &tb), &fused);
      if (!rc && !fused)
        rc = td_add_launch<MODE_STAGE>(c, x->prec, cf, ins[j], c->k[j], outs[j], cs[j], nullptr, nullptr, 0.0);
    }
    const double2 *cf = c->s + (size_t)3 * st;
    if (!rc)
      rc = qh_td_sliced_stage(G, A, x->prec, MODE_FINAL, c->y[0], nullptr, nullptr, 0.0, x->d, c->k[0], c->k[1],
                              c->k[2], x->d, h, lo, td_fusable(c, cf, &tb), &fused);
    if (!rc && !fused)
      rc = td_add_launch<MODE_FINAL>(c, x->prec, cf, c->y[0], nullptr, nullptr, 0.0, x->d, lo, h);
    return rc;
  }
  for (int j = 0; j < 3 && !rc; j++)
    rc = td_launch<MODE_STAGE>(A, c, x->prec, c->s + (size_t)j * st, ins[j], c->k[j], outs[j], cs[j], x->d, lo,
                               nullptr, nullptr, nullptr, 0.0);
  if (!rc)
    rc = td_launch<MODE_FINAL>(A, c, x->prec, c->s + (size_t)3 * st, c->y[0], nullptr, nullptr, 0.0, x->d, lo,
                               c->k[0], c->k[1], c->k[2], h);
  return rc;
}

extern "C" int qh_rk3bs_td_step(qh_mat *A, int ntd, qh_mat *const *Ak, const double *s4_re_im, const qh_vec *x,
                                qh_vec *xn, qh_vec *k1, qh_vec *k2, qh_vec *k3, qh_vec *k4, qh_vec *y, double h,
                                double atol, double rtol, int have_k1, double *err) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && x && (ntd == 0 || (Ak && s4_re_im)), QH_EINVAL, "qh_rk3bs_td_step: null argument");
  QH_REQUIRE(x->n == A->rows && x->prec == QH_FP64, QH_EINVAL,
             "qh_rk3bs_td_step: FP64 state of the matrix dimension required");
  qh_td_cache *c = nullptr;
  rc = td_cache_get(A, ntd, Ak, &c);
  if (rc) return rc;
  if (g_td_sliced && c->soff && g_td_lanes == 0) {
    // the four stage tables in one upload, the entries summed in A's stage kernels
    const int st = c->stride();
    std::vector<double2> hs((size_t)4 * st);
    for (int j = 0; j < 4; j++) td_table(c, ntd, s4_re_im + 2 * (size_t)j * ntd, hs.data() + (size_t)j * st);
    QH_HIP(hipMemcpyAsync(c->s, hs.data(), sizeof(double2) * 4 * st, hipMemcpyHostToDevice, qh_stream()));
    qh_td_ell t4[4];
    for (int j = 0; j < 4; j++) t4[j] = qh_td_ell{c->soff, c->sw, c->ecol, c->eidx, c->s + (size_t)j * st};
    int done = 0;
    rc = qh_rk3bs_td_sliced(A, x, xn, k1, k2, k3, k4, y, h, atol, rtol, have_k1, err, t4, &done);
    if (rc || done) return rc;
  }
  // four RHS evaluations and the combinations
  const double *s0 = s4_re_im, *s1 = s4_re_im + 2 * (size_t)ntd, *s2 = s4_re_im + 4 * (size_t)ntd,
               *s3 = s4_re_im + 6 * (size_t)ntd;
  if (!have_k1) rc = qh_spmv_td(A, ntd, Ak, s0, x, k1);
  if (!rc) rc = qh_axpy_to(y, x, 0.5 * h, 0.0, k1);
  if (!rc) rc = qh_spmv_td(A, ntd, Ak, s1, y, k2);
  if (!rc) rc = qh_axpy_to(y, x, 0.75 * h, 0.0, k2);
  if (!rc) rc = qh_spmv_td(A, ntd, Ak, s2, y, k3);
  if (!rc) {
    const double cf[8] = {1.0, 0.0, h * 2.0 / 9.0, 0.0, h / 3.0, 0.0, h * 4.0 / 9.0, 0.0};
    qh_vec *v[4] = {const_cast<qh_vec *>(x), k1, k2, k3};
    rc = qh_lincomb(xn, 4, cf, v);
  }
  if (!rc) rc = qh_spmv_td(A, ntd, Ak, s3, xn, k4);
  if (!rc) rc = qh_rk3bs_err(x, xn, k1, k2, k3, k4, h, atol, rtol, err);
  return rc;
}

extern "C" int qh_rk4_td_steps(qh_mat *A, int ntd, qh_mat *const *Ak, const double *s_re_im, int64_t nsteps,
                               qh_vec *x, double h) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(A && x && (ntd == 0 || (Ak && s_re_im)), QH_EINVAL, "qh_rk4_td_steps: null argument");
  QH_REQUIRE(x->n == A->rows && nsteps >= 0, QH_EINVAL, "qh_rk4_td_steps: state of the matrix dimension required");
  if (nsteps == 0) return QH_OK;
  qh_td_cache *c = nullptr;
  rc = td_cache_get(A, ntd, Ak, &c);
  if (rc) return rc;
  const int64_t n = A->rows;
  if (n > QH_TD_SMALL_ROWS || A->lanes_forced || !c->dict()) {
    // step by step
    for (int64_t j = 0; j < nsteps && !rc; j++)
      rc = qh_rk4_td_step(A, ntd, Ak, s_re_im + 8 * (size_t)ntd * j, x, h);
    return rc;
  }
  // the stage tables of every step in one upload, then one workgroup
  const int st = c->stride();
  const size_t need = (size_t)4 * st * nsteps;
  if (c->steps_cap < need) {
    qh_free_ptr(c->steps_cf);
    c->steps_cf = nullptr;
    c->steps_cap = 0;
    rc = qh_alloc_t(&c->steps_cf, need);
    if (rc) return rc;
    c->steps_cap = need;
  }
  std::vector<double2> hs(need);
  for (int64_t j = 0; j < nsteps; j++)
    for (int q = 0; q < 4; q++)
      td_table(c, ntd, s_re_im + 2 * (size_t)ntd * (4 * j + q), hs.data() + (size_t)st * (4 * j + q));
  hipStream_t sm = qh_stream();
  QH_HIP(hipMemcpyAsync(c->steps_cf, hs.data(), sizeof(double2) * need, hipMemcpyHostToDevice, sm));
  const int bt = (int)std::min<int64_t>(1024, ((n + 63) / 64) * 64);
  const size_t lds = sizeof(double2) * 5 * (size_t)n;
  if (x->prec == QH_FP32)
    hipLaunchKernelGGL((k_rk4_td_small<true>), dim3(1), dim3(bt), lds, sm, (int32_t)n, A->rowptr, A->col, A->val,
                       c->trp, c->tcol, c->tidx, c->steps_cf, st, x->d, x->lo, h, nsteps);
  else
    hipLaunchKernelGGL((k_rk4_td_small<false>), dim3(1), dim3(bt), lds, sm, (int32_t)n, A->rowptr, A->col, A->val,
                       c->trp, c->tcol, c->tidx, c->steps_cf, st, x->d, nullptr, h, nsteps);
  QH_HIP(hipGetLastError());
  // the tables are read by the kernel: keep the host copy alive until it ends
  QH_HIP(hipStreamSynchronize(sm));
  return QH_OK;
}
