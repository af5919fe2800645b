// qh_assemble.hip -- construct_matrix on the device, straight into CSR.
//
// Reference semantics (qsystem.c:593-676, kron.c:638-1036): for every term and
// every iteration row i the reference calls MatSetValue(ADD_VALUES) for up to
// three entries; PETSc keeps each row sorted, sums duplicates in insertion
// order, keeps explicit zeros, and construct_matrix finally adds 0 to every
// diagonal.  The transposed scatter of the G^T x I part (kron.c:721-725) and
// the diagonal L^dag L parts (kron.c:1012-1024) write into rows other than i.
//
// Design: every QuaC operator chain is a partial bijection of basis indices, so
// each target row r can *gather* its own contributions instead of having them
// scattered to it: the I x G and G* x G parts come from iteration r (forward
// chain), the G^T x I and diagonal parts from the unique iteration
// i = chain^{-1}(r) (inverse chain).  Ordering the <= 3 contributions of a term
// by (iteration, part) reproduces PETSc's insertion order exactly, so one
// thread per row rebuilds the row with PETSc's own insert/merge rule -- no
// global sort, no atomics, deterministic, bit-identical values.
//
// Passes: count (upper bound per row) -> scan -> fill (sorted insert/merge in a
// per-row scratch segment) -> scan of exact lengths -> compact (+ -i scale,
// + complex64 copy).
#include <hipcub/hipcub.hpp>
#include <vector>

#include "qh_common.h"

#pragma clang fp contract(off)

namespace {

struct ChainG;
struct Ctx {
  const qh_term *terms;
  int nterms;
  const qh_op *ops;
  const qh_level *tab;
  uint32_t tot;
  int conj_a;
  __device__ inline ChainG chain(int op0) const;
};

__device__ inline cxd cmul(cxd a, cxd b) { return cxd_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ inline cxd cconj(cxd a) { return cxd_make(a.re, -a.im); }

enum { M_IG = 0, M_GI = 1, M_GG = 2 };

// _get_val_j_ops (kron.c:161-207) with the per-op rule of kron.c:294-538:
// j = i + (next(sub) - sub) * n_after with sub = i / n_after % levels.  Every
// index and stride is below the row count (< 2^32), so the digit extraction
// uses 32-bit unsigned division (64-bit division is a long emulated sequence on
// the GPU).
__device__ inline uint32_t digit(uint32_t x, uint32_t stride, uint32_t levels) { return (x / stride) % levels; }

__device__ bool chain_fwd(const Ctx &c, const qh_op *ops, int n, uint32_t i, int mode, uint32_t &j, cxd &v) {
  uint32_t cur = i;
  cxd val = cxd_make(1.0, 0.0);
  for (int k = 0; k < n; k++) {
    const qh_op op = ops[k];
    const uint32_t L = (uint32_t)op.levels, s = (uint32_t)op.stride;
    if (mode == M_GG) {
      uint32_t i1 = cur / c.tot, i2 = cur - i1 * c.tot;
      const uint32_t s1 = digit(i1, s, L), s2 = digit(i2, s, L);
      const qh_level e1 = c.tab[op.tab + s1], e2 = c.tab[op.tab + s2];
      if (e1.next < 0 || e2.next < 0) return false;
      i1 = (uint32_t)((int64_t)i1 + ((int64_t)e1.next - s1) * s);
      i2 = (uint32_t)((int64_t)i2 + ((int64_t)e2.next - s2) * s);
      cur = i1 * c.tot + i2;
      cxd t = cmul(cxd_make(e1.re, -e1.im), cxd_make(e2.re, e2.im));
      val = cmul(t, val);
    } else {
      const uint32_t st = (mode == M_GI) ? s * c.tot : s;
      const uint32_t sub = digit(cur, st, L);
      const qh_level e = c.tab[op.tab + sub];
      if (e.next < 0) return false;
      cur = (uint32_t)((int64_t)cur + ((int64_t)e.next - sub) * st);
      cxd ev = cxd_make(e.re, mode == M_GI ? -e.im : e.im);
      val = cmul(ev, val);
    }
  }
  j = cur;
  v = val;
  return true;
}

// inverse of the forward chain (IG or GI): the unique i with chain(i) = r
__device__ bool chain_inv(const Ctx &c, const qh_op *ops, int n, uint32_t r, int mode, uint32_t &i) {
  uint32_t cur = r;
  for (int k = n - 1; k >= 0; k--) {
    const qh_op op = ops[k];
    const uint32_t L = (uint32_t)op.levels;
    const uint32_t st = (mode == M_GI) ? (uint32_t)op.stride * c.tot : (uint32_t)op.stride;
    const uint32_t sub = digit(cur, st, L);
    const int32_t p = c.tab[op.tab + sub].prev;
    if (p < 0) return false;
    cur = (uint32_t)((int64_t)cur + ((int64_t)p - sub) * st);
  }
  i = cur;
  return true;
}

struct Contrib {
  uint64_t key;  // iteration * 4 + part: PETSc insertion order within a term
  uint32_t col;
  cxd v;
};

// the chain rules of a context with plain division and tables in global memory
struct ChainG {
  const Ctx &c;
  const qh_op *ops;
  __device__ bool fwd(int n, uint32_t i, int mode, uint32_t &j, cxd &v) const { return chain_fwd(c, ops, n, i, mode, j, v); }
  __device__ bool inv(int n, uint32_t r, int mode, uint32_t &i) const { return chain_inv(c, ops, n, r, mode, i); }
};

__device__ inline ChainG Ctx::chain(int op0) const { return ChainG{*this, ops + op0}; }

// compare-exchange by key, field by field (a select of whole structs would go
// through memory)
__device__ inline void cswap(Contrib &a, Contrib &b) {
  const bool sw = a.key > b.key;
  const uint64_t k0 = sw ? b.key : a.key, k1 = sw ? a.key : b.key;
  const uint32_t c0 = sw ? b.col : a.col, c1 = sw ? a.col : b.col;
  const double r0 = sw ? b.v.re : a.v.re, r1 = sw ? a.v.re : b.v.re;
  const double i0 = sw ? b.v.im : a.v.im, i1 = sw ? a.v.im : b.v.im;
  a.key = k0;
  a.col = c0;
  a.v = cxd_make(r0, i0);
  b.key = k1;
  b.col = c1;
  b.v = cxd_make(r1, i1);
}

// The contributions of one term to target row r, passed to emit(col, value) in
// PETSc's insertion order (by iteration, then part): at most three, held in
// three fixed slots and ordered by a three-element sorting network, so that
// everything stays in registers.
template <class CTX, class F>
__device__ void term_emit(const CTX &c, const qh_term &t, uint32_t r, F &&emit) {
  const auto ch = c.chain(t.op0);
  const int n = t.nops;
  cxd a = t.zero_a ? cxd_make(0.0, 0.0) : cxd_make(t.a_re, t.a_im);
  const cxd I = cxd_make(0.0, 1.0);
  const uint64_t NONE = ~(uint64_t)0;
  Contrib s0 = {NONE, 0u, cxd_make(0.0, 0.0)}, s1 = s0, s2 = s0;
  uint32_t j, i;
  cxd v;
  switch (t.kind) {
    case QH_HAM_ONLY:  // kron.c:638-657: A[i][j] += a*val
      if (ch.fwd(n, r, M_IG, j, v)) s0 = {(uint64_t)r * 4, j, cmul(a, v)};
      break;
    case QH_LIN_MCWF:  // kron.c:664-700: A[i][j] += -i*a/2*val (ops list already doubled)
      if (ch.fwd(n, r, M_IG, j, v)) {
        cxd sc = cmul(cxd_make(-0.0, -1.0), a);
        sc = cxd_make(sc.re / 2, sc.im / 2);
        s0 = {(uint64_t)r * 4, j, cmul(sc, v)};
      }
      break;
    case QH_HAM_DM:  // kron.c:702-729
      if (ch.fwd(n, r, M_IG, j, v)) {
        cxd na = cxd_make(-a.re, -a.im);
        s0 = {(uint64_t)r * 4 + 0, j, cmul(cmul(na, I), v)};
      }
      if (ch.inv(n, r, M_GI, i) && ch.fwd(n, i, M_GI, j, v) && j == r) {
        cxd aa = c.conj_a ? cconj(a) : a;
        s1 = {(uint64_t)i * 4 + 1, i, cmul(cmul(aa, I), cconj(v))};
      }
      break;
    case QH_LIN_DM: {  // kron.c:995-1036
      cxd h = cxd_make(-0.5 * a.re, -0.5 * a.im);
      if (ch.inv(n, r, M_IG, i) && ch.fwd(n, i, M_IG, j, v) && j == r)
        s0 = {(uint64_t)i * 4 + 0, r, cmul(cmul(h, cconj(v)), v)};
      if (ch.inv(n, r, M_GI, i) && ch.fwd(n, i, M_GI, j, v) && j == r)
        s1 = {(uint64_t)i * 4 + 1, r, cmul(h, cconj(cmul(cconj(v), v)))};
      if (ch.fwd(n, r, M_GG, j, v)) s2 = {(uint64_t)r * 4 + 2, j, cmul(a, v)};
      break;
    }
    default:
      break;
  }
  cswap(s0, s1);
  cswap(s1, s2);
  cswap(s0, s1);
  if (s0.key != NONE) emit(s0.col, s0.v);
  if (s1.key != NONE) emit(s1.col, s1.v);
  if (s2.key != NONE) emit(s2.col, s2.v);
}

__global__ void __launch_bounds__(256) k_count(Ctx c, uint32_t rows, int64_t *cnt) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  int64_t n = 1;  // the explicit diagonal zero (qsystem.c:656-662)
  for (int t = 0; t < c.nterms; t++) term_emit(c, c.terms[t], r, [&](uint32_t, cxd) { n++; });
  cnt[r] = n;
}

// MatSetValues_SeqAIJ rule: an existing entry is added to, a new one is set.
__device__ inline void seg_insert(uint32_t *col, cxd *val, int &len, uint32_t cidx, cxd v) {
  int p = len;
  while (p > 0 && col[p - 1] > cidx) p--;
  if (p > 0 && col[p - 1] == cidx) {
    val[p - 1] = cxd_make(val[p - 1].re + v.re, val[p - 1].im + v.im);
    return;
  }
  for (int q = len; q > p; q--) {
    col[q] = col[q - 1];
    val[q] = val[q - 1];
  }
  col[p] = cidx;
  val[p] = v;
  len++;
}

__global__ void __launch_bounds__(256) k_fill(Ctx c, uint32_t rows, const int64_t *off, uint32_t *tcol, cxd *tval,
                                              int64_t *len_out) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  uint32_t *col = tcol + off[r];
  cxd *val = tval + off[r];
  int len = 0;
  for (int t = 0; t < c.nterms; t++)
    term_emit(c, c.terms[t], r, [&](uint32_t cc, cxd v) { seg_insert(col, val, len, cc, v); });
  seg_insert(col, val, len, r, cxd_make(0.0, 0.0));
  len_out[r] = len;
}

__global__ void __launch_bounds__(256) k_compact(uint32_t rows, const int64_t *off, const int64_t *len,
                                                 const int64_t *rowstart, const uint32_t *tcol, const cxd *tval,
                                                 int32_t *rowptr, int32_t *colout, double2 *valout, float2 *val32,
                                                 int scale) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  int64_t s = off[r], d = rowstart[r], n = len[r];
  rowptr[r] = (int32_t)d;
  if (r == rows - 1) rowptr[rows] = (int32_t)(d + n);
  for (int64_t k = 0; k < n; k++) {
    cxd v = tval[s + k];
    if (scale) v = cmul(cxd_make(-0.0, -1.0), v);  // MatScale(A,-i), qsystem.c:670-673
    colout[d + k] = (int32_t)tcol[s + k];
    valout[d + k] = make_double2(v.re, v.im);
    if (val32) val32[d + k] = make_float2((float)v.re, (float)v.im);
  }
}

// ------------------------------------------------------------ fast path
// Division of a 32-bit index by a fixed divisor d with a precomputed multiplier
// (Granlund & Montgomery 1994): with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (mulhi(n, m) + n) >> l for every
// n < 2^32 (the sum taken in 64 bits).
struct FDiv {
  uint32_t m, sh, d;
};
FDiv fdiv_make(uint32_t d) {
  uint32_t sh = 0;
  while (sh < 32 && ((uint64_t)1 << sh) < d) sh++;
  const uint64_t m = (((uint64_t)1 << 32) * (((uint64_t)1 << sh) - d)) / d + 1;
  return FDiv{(uint32_t)m, sh, d};
}
__device__ inline uint32_t fdiv(uint32_t n, const FDiv &f) {
  return (uint32_t)(((uint64_t)__umulhi(n, f.m) + n) >> f.sh);
}
// (x / a) % L
__device__ inline uint32_t fdigit(uint32_t x, const FDiv &a, const FDiv &L) {
  const uint32_t q = fdiv(x, a);
  return q - L.d * fdiv(q, L);
}

// an operator with its divisors: the stride, the stride times tot (the column
// digit of the transposed part), and the level count
struct OpF {
  FDiv s, sg, L;
  int32_t tab;
  uint32_t pad[2];  // 48 bytes: the level table after the operators stays 16-byte aligned
};
static_assert(sizeof(OpF) == 48, "OpF size");

struct ChainF;
struct CtxF {
  const qh_term *terms;
  int nterms;
  const OpF *ops;      // in LDS
  const qh_level *tab; // in LDS
  uint32_t tot;
  FDiv ftot;
  int conj_a;
  __device__ inline ChainF chain(int op0) const;
};

// chain_fwd / chain_inv of the plain context, with the divisions replaced by
// multiplier divisions and the tables read from LDS
struct ChainF {
  const CtxF &c;
  const OpF *ops;
  __device__ bool fwd(int n, uint32_t i, int mode, uint32_t &j, cxd &v) const {
    uint32_t cur = i;
    cxd val = cxd_make(1.0, 0.0);
    for (int k = 0; k < n; k++) {
      const OpF &op = ops[k];
      const uint32_t s = op.s.d;
      if (mode == M_GG) {
        uint32_t i1 = fdiv(cur, c.ftot), i2 = cur - i1 * c.tot;
        const uint32_t s1 = fdigit(i1, op.s, op.L), s2 = fdigit(i2, op.s, op.L);
        const qh_level e1 = c.tab[op.tab + s1], e2 = c.tab[op.tab + s2];
        if (e1.next < 0 || e2.next < 0) return false;
        i1 = (uint32_t)((int64_t)i1 + ((int64_t)e1.next - s1) * s);
        i2 = (uint32_t)((int64_t)i2 + ((int64_t)e2.next - s2) * s);
        cur = i1 * c.tot + i2;
        cxd t = cmul(cxd_make(e1.re, -e1.im), cxd_make(e2.re, e2.im));
        val = cmul(t, val);
      } else {
        const FDiv &st = (mode == M_GI) ? op.sg : op.s;
        const uint32_t sub = fdigit(cur, st, op.L);
        const qh_level e = c.tab[op.tab + sub];
        if (e.next < 0) return false;
        cur = (uint32_t)((int64_t)cur + ((int64_t)e.next - sub) * st.d);
        cxd ev = cxd_make(e.re, mode == M_GI ? -e.im : e.im);
        val = cmul(ev, val);
      }
    }
    j = cur;
    v = val;
    return true;
  }
  __device__ bool inv(int n, uint32_t r, int mode, uint32_t &i) const {
    uint32_t cur = r;
    for (int k = n - 1; k >= 0; k--) {
      const OpF &op = ops[k];
      const FDiv &st = (mode == M_GI) ? op.sg : op.s;
      const uint32_t sub = fdigit(cur, st, op.L);
      const int32_t p = c.tab[op.tab + sub].prev;
      if (p < 0) return false;
      cur = (uint32_t)((int64_t)cur + ((int64_t)p - sub) * st.d);
    }
    i = cur;
    return true;
  }
};
__device__ inline ChainF CtxF::chain(int op0) const { return ChainF{*this, ops + op0}; }

// copy the operators and level tables into LDS; returns the context on them
__device__ inline CtxF ctx_to_lds(const CtxF &g, const OpF *gops, int nops, const qh_level *gtab, int nlev) {
  extern __shared__ __align__(16) char smem[];
  OpF *sops = (OpF *)smem;
  qh_level *stab = (qh_level *)(smem + sizeof(OpF) * nops);
  for (int i = threadIdx.x; i < nops; i += blockDim.x) sops[i] = gops[i];
  for (int i = threadIdx.x; i < nlev; i += blockDim.x) stab[i] = gtab[i];
  __syncthreads();
  CtxF c = g;
  c.ops = sops;
  c.tab = stab;
  return c;
}

// Exact row lengths: the distinct columns of a row collected in KC registers
// (fully unrolled compares, no indexed register access).  A row with more than
// KC distinct columns sets *ovf (the caller then takes the general path);
// *maxlen gets the longest row.
template <int KC>
__global__ void __launch_bounds__(256) k_count_exact(CtxF g, const OpF *gops, int nops, const qh_level *gtab,
                                                     int nlev, uint32_t rows, int64_t *cnt, int *ovf, int *maxlen) {
  const CtxF c = ctx_to_lds(g, gops, nops, gtab, nlev);
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  int n = 0;
  bool over = false;
  if (r < rows) {
    uint32_t kc[KC];
#pragma unroll
    for (int q = 0; q < KC; q++) kc[q] = 0xFFFFFFFFu;
    auto put = [&](uint32_t col) {
      bool found = false;
#pragma unroll
      for (int q = 0; q < KC; q++) found |= (kc[q] == col);
      if (found) return;
      if (n == KC) {
        over = true;
        return;
      }
#pragma unroll
      for (int q = 0; q < KC; q++)
        if (q == n) kc[q] = col;
      n++;
    };
    for (int t = 0; t < c.nterms; t++) term_emit(c, c.terms[t], r, [&](uint32_t cc, cxd) { put(cc); });
    put(r);  // the explicit diagonal zero (qsystem.c:656-662)
    cnt[r] = n;
  }
  if (over) atomicOr(ovf, 1);
  // block maximum of the lengths, one atomic per block
  __shared__ int bm[4];
  int w = n;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) w = max(w, __shfl_xor(w, off, 64));
  if ((threadIdx.x & 63) == 0) bm[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(maxlen, max(max(bm[0], bm[1]), max(bm[2], bm[3])));
}

// The row with PETSc's insert/merge rule (a new column is set, an existing one
// added to, in insertion order) in KC register slots, then written in column
// order straight into the final CSR (-i scale and the complex64 copy applied).
template <int KC>
__global__ void __launch_bounds__(256) k_fill_exact(CtxF g, const OpF *gops, int nops, const qh_level *gtab,
                                                    int nlev, uint32_t rows, const int64_t *start, int32_t *rowptr,
                                                    int32_t *colout, double2 *valout, float2 *val32, int scale) {
  const CtxF c = ctx_to_lds(g, gops, nops, gtab, nlev);
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  uint32_t kc[KC];
  double vr[KC], vi[KC];
#pragma unroll
  for (int q = 0; q < KC; q++) {
    kc[q] = 0xFFFFFFFFu;
    vr[q] = vi[q] = 0.0;
  }
  int n = 0;
  auto put = [&](uint32_t col, cxd v) {
    bool found = false;
#pragma unroll
    for (int q = 0; q < KC; q++)
      if (kc[q] == col) {
        vr[q] = vr[q] + v.re;
        vi[q] = vi[q] + v.im;
        found = true;
      }
    if (found || n == KC) return;  // a longer row was excluded by the count pass
#pragma unroll
    for (int q = 0; q < KC; q++)
      if (q == n) {
        kc[q] = col;
        vr[q] = v.re;
        vi[q] = v.im;
      }
    n++;
  };
  for (int t = 0; t < c.nterms; t++) term_emit(c, c.terms[t], r, [&](uint32_t cc, cxd v) { put(cc, v); });
  put(r, cxd_make(0.0, 0.0));
  const int64_t d = start[r];
  rowptr[r] = (int32_t)d;
  if (r == rows - 1) rowptr[rows] = (int32_t)(d + n);
#pragma unroll
  for (int q = 0; q < KC; q++) {
    if (q < n) {
      int rank = 0;
#pragma unroll
      for (int u = 0; u < KC; u++) rank += (kc[u] < kc[q]) ? 1 : 0;
      cxd v = cxd_make(vr[q], vi[q]);
      if (scale) v = cmul(cxd_make(-0.0, -1.0), v);  // MatScale(A,-i), qsystem.c:670-673
      colout[d + rank] = (int32_t)kc[q];
      valout[d + rank] = make_double2(v.re, v.im);
      if (val32) val32[d + rank] = make_float2((float)v.re, (float)v.im);
    }
  }
}

int exclusive_scan(const int64_t *in, int64_t *out, int64_t n, int64_t *total) {
  void *tmp = nullptr;
  size_t tmpb = 0;
  QH_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmpb, in, out, (int)n, qh_stream()));
  int rc = qh_alloc(&tmp, tmpb);
  if (rc) return rc;
  QH_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmpb, in, out, (int)n, qh_stream()));
  int64_t last_in = 0, last_out = 0;
  QH_HIP(hipMemcpyAsync(&last_in, in + n - 1, 8, hipMemcpyDeviceToHost, qh_stream()));
  QH_HIP(hipMemcpyAsync(&last_out, out + n - 1, 8, hipMemcpyDeviceToHost, qh_stream()));
  QH_HIP(hipStreamSynchronize(qh_stream()));
  qh_free_ptr(tmp);
  *total = last_in + last_out;
  return QH_OK;
}

}  // namespace

extern "C" int qh_assemble(int64_t tot, int64_t dim, int nterms, const qh_term *terms, int nops, const qh_op *ops,
                           int nlev, const qh_level *levels, int flags, qh_mat **out) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(out && tot > 0 && dim > 0, QH_EINVAL, "qh_assemble: bad arguments");
  QH_REQUIRE(dim < (int64_t)INT32_MAX && tot * tot <= (int64_t)INT32_MAX * 2, QH_ERANGE,
             "qh_assemble: dimension %lld exceeds the 32-bit CSR index range", (long long)dim);
  QH_REQUIRE(nterms >= 0 && nops >= 0 && nlev >= 0, QH_EINVAL, "qh_assemble: negative counts");
  for (int t = 0; t < nterms; t++) {
    QH_REQUIRE(terms[t].op0 >= 0 && terms[t].nops >= 0 && terms[t].op0 + terms[t].nops <= nops, QH_EINVAL,
               "qh_assemble: term %d op range out of bounds", t);
    QH_REQUIRE(terms[t].kind >= QH_HAM_DM && terms[t].kind <= QH_LIN_MCWF, QH_EINVAL, "qh_assemble: bad kind");
  }
  for (int k = 0; k < nops; k++) {
    QH_REQUIRE(ops[k].levels > 0 && ops[k].stride > 0 && ops[k].tab >= 0 && ops[k].tab + ops[k].levels <= nlev,
               QH_EINVAL, "qh_assemble: op %d table out of bounds", k);
    QH_REQUIRE(tot % (ops[k].stride * ops[k].levels) == 0, QH_EINVAL, "qh_assemble: op %d stride inconsistent", k);
  }
  for (int l = 0; l < nlev; l++) {
    // next/prev must stay inside the op's level range; checked against levels of the owner below
    QH_REQUIRE(levels[l].next >= -1 && levels[l].prev >= -1, QH_EINVAL, "qh_assemble: bad level entry");
  }
  for (int k = 0; k < nops; k++)
    for (int l = 0; l < ops[k].levels; l++) {
      const qh_level &e = levels[ops[k].tab + l];
      QH_REQUIRE(e.next < ops[k].levels && e.prev < ops[k].levels, QH_EINVAL,
                 "qh_assemble: op %d level map leaves the subsystem", k);
    }

  hipStream_t s = qh_stream();
  const uint32_t rows = (uint32_t)dim;
  qh_term *d_terms = nullptr;
  qh_op *d_ops = nullptr;
  qh_level *d_lev = nullptr;
  int64_t *cnt = nullptr, *off = nullptr, *len = nullptr, *start = nullptr;
  uint32_t *tcol = nullptr;
  cxd *tval = nullptr;
  qh_mat *m = nullptr;
  int64_t total = 0, nnz = 0;
  Ctx c;
  OpF *d_opf = nullptr;
  int *d_flags = nullptr;

#define QH_TRY(x)          \
  do {                     \
    rc = (x);              \
    if (rc) goto cleanup;  \
  } while (0)
#define QH_TRYHIP(x)                                                                  \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      qh_set_error("qh_assemble: %s: %s", #x, hipGetErrorString(e_));                 \
      rc = QH_EHIP;                                                                   \
      goto cleanup;                                                                   \
    }                                                                                 \
  } while (0)

  QH_TRY(qh_alloc_t(&d_terms, nterms > 0 ? nterms : 1));
  QH_TRY(qh_alloc_t(&d_ops, nops > 0 ? nops : 1));
  QH_TRY(qh_alloc_t(&d_lev, nlev > 0 ? nlev : 1));
  if (nterms) QH_TRYHIP(hipMemcpyAsync(d_terms, terms, sizeof(qh_term) * nterms, hipMemcpyHostToDevice, s));
  if (nops) QH_TRYHIP(hipMemcpyAsync(d_ops, ops, sizeof(qh_op) * nops, hipMemcpyHostToDevice, s));
  if (nlev) QH_TRYHIP(hipMemcpyAsync(d_lev, levels, sizeof(qh_level) * nlev, hipMemcpyHostToDevice, s));
  c.terms = d_terms;
  c.nterms = nterms;
  c.ops = d_ops;
  c.tab = d_lev;
  c.tot = (uint32_t)tot;
  c.conj_a = (flags & QH_ASM_CONJ_A) ? 1 : 0;

  QH_TRY(qh_alloc_t(&cnt, rows));
  QH_TRY(qh_alloc_t(&off, rows));
  {
    // fast path: exact row lengths with the distinct columns in registers, rows
    // written straight into the final CSR; taken when the operator tables fit
    // in LDS and no row has more than 32 distinct columns
    const size_t lds = sizeof(OpF) * (size_t)nops + sizeof(qh_level) * (size_t)nlev;
    if (lds <= 60 * 1024) {
      std::vector<OpF> hf(nops > 0 ? nops : 1);
      for (int k = 0; k < nops; k++) {
        hf[k].s = fdiv_make((uint32_t)ops[k].stride);
        hf[k].sg = fdiv_make((uint32_t)(ops[k].stride * tot));
        hf[k].L = fdiv_make((uint32_t)ops[k].levels);
        hf[k].tab = ops[k].tab;
        hf[k].pad[0] = hf[k].pad[1] = 0;
      }
      QH_TRY(qh_alloc_t(&d_opf, hf.size()));
      QH_TRY(qh_alloc_t(&d_flags, 2));
      QH_TRYHIP(hipMemcpyAsync(d_opf, hf.data(), sizeof(OpF) * hf.size(), hipMemcpyHostToDevice, s));
      QH_TRYHIP(hipMemsetAsync(d_flags, 0, 2 * sizeof(int), s));
      CtxF cf;
      cf.terms = d_terms;
      cf.nterms = nterms;
      cf.ops = nullptr;
      cf.tab = nullptr;
      cf.tot = (uint32_t)tot;
      cf.ftot = fdiv_make((uint32_t)tot);
      cf.conj_a = c.conj_a;
      hipLaunchKernelGGL(k_count_exact<32>, dim3(qh_blocks(rows, 256)), dim3(256), lds, s, cf, d_opf, nops, d_lev,
                         nlev, rows, cnt, d_flags, d_flags + 1);
      QH_TRYHIP(hipGetLastError());
      int hf2[2] = {0, 0};
      QH_TRYHIP(hipMemcpyAsync(hf2, d_flags, sizeof(hf2), hipMemcpyDeviceToHost, s));
      QH_TRYHIP(hipStreamSynchronize(s));
      if (!hf2[0]) {
        QH_TRY(exclusive_scan(cnt, off, rows, &nnz));
        if (nnz >= (int64_t)INT32_MAX) {
          qh_set_error("qh_assemble: nnz %lld exceeds the 32-bit CSR range", (long long)nnz);
          rc = QH_ERANGE;
          goto cleanup;
        }
        m = new qh_mat();
        m->rows = dim;
        m->nnz = nnz;
        QH_TRY(qh_alloc_t(&m->rowptr, (size_t)rows + 1));
        QH_TRY(qh_alloc_t(&m->col, (size_t)nnz));
        QH_TRY(qh_alloc_t(&m->val, (size_t)nnz));
        if (flags & QH_ASM_FP32) QH_TRY(qh_alloc_t(&m->val32, (size_t)nnz));
        const int sc = (flags & QH_ASM_SCALE_MINUS_I) ? 1 : 0;
        const dim3 g(qh_blocks(rows, 256));
#define QH_FILLX(KC)                                                                                              \
  hipLaunchKernelGGL(k_fill_exact<KC>, g, dim3(256), lds, s, cf, d_opf, nops, d_lev, nlev, rows, off, m->rowptr, \
                     m->col, m->val, m->val32, sc)
        if (hf2[1] <= 8) QH_FILLX(8);
        else if (hf2[1] <= 16) QH_FILLX(16);
        else QH_FILLX(32);
#undef QH_FILLX
        QH_TRYHIP(hipGetLastError());
        QH_TRYHIP(hipStreamSynchronize(s));
        *out = m;
        m = nullptr;
        goto cleanup;
      }
    }
  }
  hipLaunchKernelGGL(k_count, dim3(qh_blocks(rows, 256)), dim3(256), 0, s, c, rows, cnt);
  QH_TRYHIP(hipGetLastError());
  QH_TRY(exclusive_scan(cnt, off, rows, &total));
  QH_TRY(qh_alloc_t(&tcol, (size_t)total));
  QH_TRY(qh_alloc_t(&tval, (size_t)total));
  QH_TRY(qh_alloc_t(&len, rows));
  hipLaunchKernelGGL(k_fill, dim3(qh_blocks(rows, 256)), dim3(256), 0, s, c, rows, off, tcol, tval, len);
  QH_TRYHIP(hipGetLastError());
  start = cnt;  // reuse
  QH_TRY(exclusive_scan(len, start, rows, &nnz));
  if (nnz >= (int64_t)INT32_MAX) {
    qh_set_error("qh_assemble: nnz %lld exceeds the 32-bit CSR range", (long long)nnz);
    rc = QH_ERANGE;
    goto cleanup;
  }
  m = new qh_mat();
  m->rows = dim;
  m->nnz = nnz;
  QH_TRY(qh_alloc_t(&m->rowptr, (size_t)rows + 1));
  QH_TRY(qh_alloc_t(&m->col, (size_t)nnz));
  QH_TRY(qh_alloc_t(&m->val, (size_t)nnz));
  if (flags & QH_ASM_FP32) QH_TRY(qh_alloc_t(&m->val32, (size_t)nnz));
  hipLaunchKernelGGL(k_compact, dim3(qh_blocks(rows, 256)), dim3(256), 0, s, rows, off, len, start, tcol, tval,
                     m->rowptr, m->col, m->val, m->val32, (flags & QH_ASM_SCALE_MINUS_I) ? 1 : 0);
  QH_TRYHIP(hipGetLastError());
  QH_TRYHIP(hipStreamSynchronize(s));
  *out = m;
  m = nullptr;
cleanup:
  if (m) qh_mat_destroy(m);
  qh_free_ptr(d_opf);
  qh_free_ptr(d_flags);
  qh_free_ptr(d_terms);
  qh_free_ptr(d_ops);
  qh_free_ptr(d_lev);
  qh_free_ptr(cnt);
  qh_free_ptr(off);
  qh_free_ptr(len);
  qh_free_ptr(tcol);
  qh_free_ptr(tval);
  return rc;
#undef QH_TRY
#undef QH_TRYHIP
}

extern "C" int qh_mat_destroy(qh_mat *m) {
  if (!m) return QH_OK;
  qh_td_cache_free(m->td_cache);
  qh_free_ptr(m->rowptr);
  qh_free_ptr(m->col);
  qh_free_ptr(m->val);
  qh_free_ptr(m->val32);
  qh_free_ptr(m->ks_moff);
  qh_free_ptr(m->ks_umid);
  qh_free_ptr(m->ks_cmid);
  qh_free_ptr(m->ks2_moff);
  qh_free_ptr(m->ks2_wmid);
  qh_free_ptr(m->ks2_smid);
  qh_free_ptr(m->ks2_umid);
  qh_free_ptr(m->ks2_emid);
  qh_free_ptr(m->ks2_cmid);
  qh_free_ptr(m->ks2_meta);
  qh_free_ptr(m->ks2_val);
  qh_free_ptr(m->ks2_sval);
  qh_free_ptr(m->ks2_sval1);
  qh_free_ptr(m->ks2_xmid);
  qh_free_ptr(m->ks2_sep);
  qh_free_ptr(m->ks2_gh);
  qh_free_ptr(m->ks_sval);
  qh_free_ptr(m->ks_sval32);
  qh_free_ptr(m->ks_meta);
  qh_free_ptr(m->ks_val);
  qh_free_ptr(m->ks_val32);
  qh_free_ptr(m->ks_meta32);
  delete m;
  return QH_OK;
}

extern "C" int qh_mat_info(const qh_mat *m, int64_t *rows, int64_t *nnz) {
  QH_REQUIRE(m, QH_EINVAL, "null matrix");
  if (rows) *rows = m->rows;
  if (nnz) *nnz = m->nnz;
  return QH_OK;
}

__global__ void k_widen_i32(const int32_t *a, int64_t *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

extern "C" int qh_mat_export(const qh_mat *m, int64_t *rowptr, int64_t *col, double *val) {
  QH_REQUIRE(m, QH_EINVAL, "null matrix");
  hipStream_t s = qh_stream();
  int64_t *tmp = nullptr;
  int64_t big = m->rows + 1 > m->nnz ? m->rows + 1 : m->nnz;
  int rc = qh_alloc_t(&tmp, (size_t)big);
  if (rc) return rc;
  if (rowptr) {
    hipLaunchKernelGGL(k_widen_i32, dim3(qh_blocks(m->rows + 1, 256)), dim3(256), 0, s, m->rowptr, tmp, m->rows + 1);
    QH_HIP(hipMemcpyAsync(rowptr, tmp, 8 * (m->rows + 1), hipMemcpyDeviceToHost, s));
    QH_HIP(hipStreamSynchronize(s));
  }
  if (col && m->nnz) {
    hipLaunchKernelGGL(k_widen_i32, dim3(qh_blocks(m->nnz, 256)), dim3(256), 0, s, m->col, tmp, m->nnz);
    QH_HIP(hipMemcpyAsync(col, tmp, 8 * m->nnz, hipMemcpyDeviceToHost, s));
    QH_HIP(hipStreamSynchronize(s));
  }
  if (val && m->nnz) QH_HIP(hipMemcpyAsync(val, m->val, 16 * m->nnz, hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  qh_free_ptr(tmp);
  return QH_OK;
}

__global__ void k_narrow_vals(const double2 *a, float2 *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

__global__ void k_narrow_reals(const double *a, float *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

extern "C" int qh_mat_make_fp32(qh_mat *m) {
  QH_REQUIRE(m, QH_EINVAL, "null matrix");
  if (!m->val32) {
    int rc = qh_alloc_t(&m->val32, (size_t)(m->nnz ? m->nnz : 1));
    if (rc) return rc;
  }
  if (m->nnz)
    hipLaunchKernelGGL(k_narrow_vals, dim3(qh_blocks(m->nnz, 256)), dim3(256), 0, qh_stream(), m->val, m->val32,
                       m->nnz);
  if (m->ks_val32 && m->ks_nval)  // refresh the sliced-layout copies too
    hipLaunchKernelGGL(k_narrow_vals, dim3(qh_blocks(m->ks_nval, 256)), dim3(256), 0, qh_stream(), m->ks_val,
                       m->ks_val32, m->ks_nval);
  if (m->ks_sval32 && m->ks_nsval)
    hipLaunchKernelGGL(k_narrow_reals, dim3(qh_blocks(m->ks_nsval, 256)), dim3(256), 0, qh_stream(), m->ks_sval,
                       m->ks_sval32, m->ks_nsval);
  QH_HIP(hipGetLastError());
  return QH_OK;
}

__global__ void k_bd_rows(const int32_t *rp, int64_t rows, int32_t row0, int32_t nz0, int32_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= rows; i += (int64_t)gridDim.x * blockDim.x)
    out[row0 + i] = rp[i] + nz0;
}
__global__ void k_bd_cols(const int32_t *c, int64_t nnz, int32_t row0, int32_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = c[i] + row0;
}

extern "C" int qh_mat_blockdiag(int n, qh_mat *const *mats, qh_mat **out) {
  QH_REQUIRE(n > 0 && mats && out, QH_EINVAL, "qh_mat_blockdiag: bad arguments");
  int64_t rows = 0, nnz = 0;
  for (int i = 0; i < n; i++) {
    QH_REQUIRE(mats[i], QH_EINVAL, "qh_mat_blockdiag: null matrix %d", i);
    rows += mats[i]->rows;
    nnz += mats[i]->nnz;
  }
  QH_REQUIRE(rows < INT32_MAX && nnz < INT32_MAX, QH_ERANGE, "qh_mat_blockdiag: too large for 32-bit CSR");
  qh_mat *m = new qh_mat();
  m->rows = rows;
  m->nnz = nnz;
  int rc = qh_alloc_t(&m->rowptr, (size_t)rows + 1);
  if (!rc) rc = qh_alloc_t(&m->col, (size_t)nnz);
  if (!rc) rc = qh_alloc_t(&m->val, (size_t)nnz);
  if (rc) {
    qh_mat_destroy(m);
    return rc;
  }
  hipStream_t s = qh_stream();
  int64_t r0 = 0, z0 = 0;
  for (int i = 0; i < n; i++) {
    const qh_mat *a = mats[i];
    hipLaunchKernelGGL(k_bd_rows, dim3(qh_blocks(a->rows + 1, 256)), dim3(256), 0, s, a->rowptr, a->rows,
                       (int32_t)r0, (int32_t)z0, m->rowptr);
    hipLaunchKernelGGL(k_bd_cols, dim3(qh_blocks(a->nnz, 256)), dim3(256), 0, s, a->col, a->nnz, (int32_t)r0,
                       m->col + z0);
    QH_HIP(hipMemcpyAsync(m->val + z0, a->val, 16 * a->nnz, hipMemcpyDeviceToDevice, s));
    r0 += a->rows;
    z0 += a->nnz;
  }
  QH_HIP(hipGetLastError());
  QH_HIP(hipStreamSynchronize(s));
  m->blk_off.resize(n + 1);
  m->blk_off[0] = 0;
  for (int i = 0; i < n; i++) m->blk_off[i + 1] = m->blk_off[i] + mats[i]->rows;
  *out = m;
  return QH_OK;
}

/* A device matrix from a host CSR: square, rows in order, strictly increasing
 * columns within a row, complex128 values interleaved.  Everything is checked
 * on the host before the upload, so the SpMV kernels never see a column out of
 * range. */
extern "C" int qh_mat_from_csr(int64_t rows, const int64_t *rowptr, const int64_t *col, const double *val,
                               qh_mat **out) {
  QH_REQUIRE(rows > 0 && rowptr && col && val && out, QH_EINVAL, "qh_mat_from_csr: bad arguments");
  const int64_t nnz = rowptr[rows];
  QH_REQUIRE(rowptr[0] == 0 && nnz > 0 && rows < INT32_MAX && nnz < INT32_MAX, QH_ERANGE,
             "qh_mat_from_csr: rowptr[0] must be 0 and 0 < nnz < 2^31");
  std::vector<int32_t> rp32((size_t)rows + 1), c32((size_t)nnz);
  for (int64_t r = 0; r < rows; r++) {
    QH_REQUIRE(rowptr[r] <= rowptr[r + 1] && rowptr[r + 1] <= nnz, QH_EINVAL, "qh_mat_from_csr: rowptr not monotone");
    rp32[r] = (int32_t)rowptr[r];
    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; e++) {
      QH_REQUIRE(col[e] >= 0 && col[e] < rows && (e == rowptr[r] || col[e] > col[e - 1]), QH_EINVAL,
                 "qh_mat_from_csr: row %lld has a column out of range or out of order", (long long)r);
      c32[e] = (int32_t)col[e];
    }
  }
  rp32[rows] = (int32_t)nnz;
  qh_mat *m = new qh_mat();
  m->rows = rows;
  m->nnz = nnz;
  int rc = qh_alloc_t(&m->rowptr, (size_t)rows + 1);
  if (!rc) rc = qh_alloc_t(&m->col, (size_t)nnz);
  if (!rc) rc = qh_alloc_t(&m->val, (size_t)nnz);
  if (rc) {
    qh_mat_destroy(m);
    return rc;
  }
  hipError_t e = hipMemcpy(m->rowptr, rp32.data(), sizeof(int32_t) * ((size_t)rows + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->col, c32.data(), sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->val, val, 16 * (size_t)nnz, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    qh_mat_destroy(m);
    QH_HIP(e);
  }
  *out = m;
  return QH_OK;
}
