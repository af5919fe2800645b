// qh_obs.hip -- observables in FP64 on the device: Tr(O rho), <psi|O|psi>,
// the trace, and all subsystem populations in one pass over the diagonal.
// qvec_utilities.c:322-392 (trace), :1197-1281 (DM expectation), :1682-2022
// (bitstring probabilities = Re diag rho).
#include <mutex>
#include <vector>

#include "qh_common.h"

namespace {

const int NB = 512;  // partial-sum blocks

// load entry i of an FP64 or FP32(hi+lo) vector as complex128
__device__ inline double2 ld(const void *d, const float2 *lo, int64_t i) {
  if (!lo) return ((const double2 *)d)[i];
  float2 h = ((const float2 *)d)[i], l = lo[i];
  return make_double2((double)h.x + (double)l.x, (double)h.y + (double)l.y);
}

// forward op chain on index i in [0, tot) (the TENSOR_IG rule of kron.c:294-538)
__device__ inline bool chain_ig(const qh_op *ops, int n, const qh_level *tab, int64_t i, int64_t &j, double2 &v) {
  int64_t cur = i;
  double vr = 1.0, vi = 0.0;
  for (int k = 0; k < n; k++) {
    const qh_op op = ops[k];
    const int64_t sub = (cur / op.stride) % op.levels;
    const qh_level e = tab[op.tab + sub];
    if (e.next < 0) return false;
    cur += (e.next - sub) * op.stride;
    const double nr = e.re * vr - e.im * vi, ni = e.re * vi + e.im * vr;
    vr = nr;
    vi = ni;
  }
  j = cur;
  v = make_double2(vr, vi);
  return true;
}

__device__ inline void block_sum2(double &a, double &b, double2 *partial) {
  __shared__ double sa[256], sb[256];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sa[threadIdx.x] += sa[threadIdx.x + o];
      sb[threadIdx.x] += sb[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(sa[0], sb[0]);
}

__global__ void __launch_bounds__(256) k_expect_dm(const void *x, const float2 *lo, int64_t tot, const qh_op *ops,
                                                   int nops, const qh_level *tab, double2 *partial) {
  double ar = 0.0, ai = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < tot; i += (int64_t)gridDim.x * 256) {
    int64_t k;
    double2 v;
    if (!chain_ig(ops, nops, tab, i, k, v)) continue;
    const double2 r = ld(x, lo, tot * i + k);  // rho[k, i]
    ar += v.x * r.x - v.y * r.y;
    ai += v.x * r.y + v.y * r.x;
  }
  block_sum2(ar, ai, partial);
}

__global__ void __launch_bounds__(256) k_expect_wf(const void *x, const float2 *lo, int64_t tot, const qh_op *ops,
                                                   int nops, const qh_level *tab, double2 *partial) {
  double ar = 0.0, ai = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < tot; i += (int64_t)gridDim.x * 256) {
    int64_t k;
    double2 v;
    if (!chain_ig(ops, nops, tab, i, k, v)) continue;
    const double2 pk = ld(x, lo, k), pi = ld(x, lo, i);
    const double tr = v.x * pk.x - v.y * pk.y, ti = v.x * pk.y + v.y * pk.x;
    ar += pi.x * tr + pi.y * ti;  // conj(psi_i) * t
    ai += pi.x * ti - pi.y * tr;
  }
  block_sum2(ar, ai, partial);
}

// ensemble of M wavefunctions (trajectory b at [b*tot, (b+1)*tot)): block b writes
// <psi_b|O|psi_b> / <psi_b|psi_b> (qvec_utilities.c:1353-1420)
__global__ void __launch_bounds__(256) k_expect_ens(const double2 *x, int64_t tot, const qh_op *ops, int nops,
                                                    const qh_level *tab, double2 *out) {
  const double2 *p = x + (int64_t)blockIdx.x * tot;
  double ar = 0.0, ai = 0.0, nn = 0.0;
  for (int64_t i = threadIdx.x; i < tot; i += 256) {
    const double2 pi = p[i];
    nn += pi.x * pi.x + pi.y * pi.y;
    int64_t k;
    double2 v;
    if (!chain_ig(ops, nops, tab, i, k, v)) continue;
    const double2 pk = p[k];
    const double tr = v.x * pk.x - v.y * pk.y, ti = v.x * pk.y + v.y * pk.x;
    ar += pi.x * tr + pi.y * ti;
    ai += pi.x * ti - pi.y * tr;
  }
  __shared__ double sa[256], sb[256], sn[256];
  sa[threadIdx.x] = ar;
  sb[threadIdx.x] = ai;
  sn[threadIdx.x] = nn;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sa[threadIdx.x] += sa[threadIdx.x + o];
      sb[threadIdx.x] += sb[threadIdx.x + o];
      sn[threadIdx.x] += sn[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = make_double2(sa[0] / sn[0], sb[0] / sn[0]);
}

__global__ void __launch_bounds__(256) k_trace(const void *x, const float2 *lo, int64_t tot, double2 *partial) {
  double ar = 0.0, ai = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < tot; i += (int64_t)gridDim.x * 256) {
    const double2 r = ld(x, lo, i * (tot + 1));
    ar += r.x;
    ai += r.y;
  }
  block_sum2(ar, ai, partial);
}

const int MAXSUB = 32;
struct SubDesc {
  int nsub;
  int64_t levels[MAXSUB], strides[MAXSUB];
};

__global__ void __launch_bounds__(256) k_pops(const void *x, const float2 *lo, int64_t tot, int dm, SubDesc d,
                                              double *partial) {
  double acc[MAXSUB];
  for (int s = 0; s < d.nsub; s++) acc[s] = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < tot; i += (int64_t)gridDim.x * 256) {
    double p;
    if (dm) {
      p = ld(x, lo, i * (tot + 1)).x;
    } else {
      const double2 a = ld(x, lo, i);
      p = a.x * a.x + a.y * a.y;
    }
    for (int s = 0; s < d.nsub; s++) acc[s] += (double)((i / d.strides[s]) % d.levels[s]) * p;
  }
  __shared__ double sh[256];
  for (int s = 0; s < d.nsub; s++) {
    sh[threadIdx.x] = acc[s];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x * MAXSUB + s] = sh[0];
    __syncthreads();
  }
}

int finish2(double2 *dpart, int nb, double *re, double *im) {
  double2 *h = nullptr;
  int rc = qh_scratch(0, nullptr, sizeof(double2) * nb, (void **)&h);
  if (rc) return rc;
  QH_HIP(hipMemcpyAsync(h, dpart, sizeof(double2) * nb, hipMemcpyDeviceToHost, qh_stream()));
  QH_HIP(hipStreamSynchronize(qh_stream()));
  double a = 0.0, b = 0.0;
  for (int i = 0; i < nb; i++) {
    a += h[i].x;
    b += h[i].y;
  }
  *re = a;
  *im = b;
  return QH_OK;
}

int blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > NB) b = NB;
  if (b < 1) b = 1;
  return (int)b;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The op chain and the NB partial sums of one reduction in the scratch buffer:
// [ops | levels | partials], each part 256-byte aligned.
int upload_chain(int nops, const qh_op *ops, int nlev, const qh_level *lev, qh_op **dops, qh_level **dlev,
                 size_t part_bytes, void **part) {
  for (int k = 0; k < nops; k++)
    QH_REQUIRE(ops[k].levels > 0 && ops[k].stride > 0 && ops[k].tab >= 0 && ops[k].tab + ops[k].levels <= nlev,
               QH_EINVAL, "observable op %d out of range", k);
  const size_t bo = align256(sizeof(qh_op) * (nops > 0 ? nops : 1)),
               bl = align256(sizeof(qh_level) * (nlev > 0 ? nlev : 1));
  char *d = nullptr;
  int rc = qh_scratch(bo + bl + part_bytes, (void **)&d, 0, nullptr);
  if (rc) return rc;
  *dops = (qh_op *)d;
  *dlev = (qh_level *)(d + bo);
  *part = d + bo + bl;
  if (nops) QH_HIP(hipMemcpyAsync(*dops, ops, sizeof(qh_op) * nops, hipMemcpyHostToDevice, qh_stream()));
  if (nlev) QH_HIP(hipMemcpyAsync(*dlev, lev, sizeof(qh_level) * nlev, hipMemcpyHostToDevice, qh_stream()));
  return QH_OK;
}

int expect_impl(int dm, const qh_vec *x, int64_t tot, int nops, const qh_op *ops, int nlev, const qh_level *lev,
                double *re, double *im) {
  QH_REQUIRE(x && re && im, QH_EINVAL, "expectation: null argument");
  QH_REQUIRE(dm ? (tot * tot == x->n) : (tot == x->n), QH_EINVAL, "expectation: vector size mismatch");
  qh_op *dops = nullptr;
  qh_level *dlev = nullptr;
  double2 *part = nullptr;
  int rc = upload_chain(nops, ops, nlev, lev, &dops, &dlev, sizeof(double2) * NB, (void **)&part);
  if (!rc) {
    const int nb = blocks_for(tot);
    if (dm)
      hipLaunchKernelGGL(k_expect_dm, dim3(nb), dim3(256), 0, qh_stream(), (const void *)x->d, x->lo, tot, dops,
                         nops, dlev, part);
    else
      hipLaunchKernelGGL(k_expect_wf, dim3(nb), dim3(256), 0, qh_stream(), (const void *)x->d, x->lo, tot, dops,
                         nops, dlev, part);
    rc = finish2(part, nb, re, im);
  }
  return rc;
}

}  // namespace

extern "C" int qh_expect_dm(const qh_vec *x, int64_t tot, int nops, const qh_op *ops, int nlev,
                            const qh_level *levels, double *re, double *im) {
  return expect_impl(1, x, tot, nops, ops, nlev, levels, re, im);
}

extern "C" int qh_expect_wf(const qh_vec *x, int64_t tot, int nops, const qh_op *ops, int nlev,
                            const qh_level *levels, double *re, double *im) {
  return expect_impl(0, x, tot, nops, ops, nlev, levels, re, im);
}

extern "C" int qh_trace(const qh_vec *x, int64_t tot, double *re, double *im) {
  QH_REQUIRE(x && tot * tot == x->n, QH_EINVAL, "qh_trace: not a density matrix of dimension %lld", (long long)tot);
  double2 *part = nullptr;
  int rc = qh_scratch(sizeof(double2) * NB, (void **)&part, 0, nullptr);
  if (rc) return rc;
  const int nb = blocks_for(tot);
  hipLaunchKernelGGL(k_trace, dim3(nb), dim3(256), 0, qh_stream(), (const void *)x->d, x->lo, tot, part);
  return finish2(part, nb, re, im);
}

extern "C" int qh_populations(const qh_vec *x, int nsub, const int64_t *levels, const int64_t *strides, int dm,
                              double *pops) {
  QH_REQUIRE(x && pops && nsub > 0 && nsub <= MAXSUB, QH_EINVAL, "qh_populations: 1..32 subsystems");
  int64_t tot = 1;
  SubDesc d;
  d.nsub = nsub;
  for (int s = 0; s < nsub; s++) {
    QH_REQUIRE(levels[s] > 0 && strides[s] > 0, QH_EINVAL, "qh_populations: bad subsystem %d", s);
    d.levels[s] = levels[s];
    d.strides[s] = strides[s];
    tot *= levels[s];
  }
  QH_REQUIRE(dm ? tot * tot == x->n : tot == x->n, QH_EINVAL, "qh_populations: vector size mismatch");
  double *part = nullptr, *h = nullptr;
  const int nb = blocks_for(tot);
  int rc = qh_scratch(sizeof(double) * NB * MAXSUB, (void **)&part, sizeof(double) * (size_t)nb * MAXSUB, (void **)&h);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pops, dim3(nb), dim3(256), 0, qh_stream(), (const void *)x->d, x->lo, tot, dm, d, part);
  QH_HIP(hipMemcpyAsync(h, part, sizeof(double) * (size_t)nb * MAXSUB, hipMemcpyDeviceToHost, qh_stream()));
  QH_HIP(hipStreamSynchronize(qh_stream()));
  for (int s = 0; s < nsub; s++) {
    double a = 0.0;
    for (int b = 0; b < nb; b++) a += h[(size_t)b * MAXSUB + s];
    pops[s] = a;
  }
  return QH_OK;
}

// ------------------------------------------------------- batched populations
extern "C" qh_vec *qh_batch_state_vec(qh_batch *b);
extern "C" int64_t qh_batch_offset(qh_batch *b, int i);

namespace {
__global__ void __launch_bounds__(256) k_batch_pops(const void *x, const float2 *lo, const int64_t *off,
                                                    const int64_t *tot, const int *nsub, const int *soff,
                                                    const int64_t *levels, const int64_t *strides, double *out) {
  const int b = blockIdx.x;
  const int64_t N = tot[b], base = off[b];
  const int ns = nsub[b], so = soff[b];
  __shared__ double sh[256];
  for (int s = 0; s < ns; s++) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += 256) {
      const double p = ld(x, lo, base + i * (N + 1)).x;
      acc += (double)((i / strides[so + s]) % levels[so + s]) * p;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[so + s] = sh[0];
    __syncthreads();
  }
}
}  // namespace

extern "C" int qh_batch_populations(qh_batch *b, const int *nsub, const int64_t *levels, const int64_t *strides,
                                    double *pops) {
  QH_REQUIRE(b && nsub && levels && strides && pops, QH_EINVAL, "qh_batch_populations: null argument");
  qh_vec *x = qh_batch_state_vec(b);
  int n = 0;
  while (qh_batch_offset(b, n) >= 0) n++;
  std::vector<int> soff(n);
  std::vector<int64_t> off(n), tot(n);
  int total = 0;
  for (int i = 0; i < n; i++) {
    soff[i] = total;
    int64_t t = 1;
    for (int s = 0; s < nsub[i]; s++) t *= levels[total + s];
    tot[i] = t;
    off[i] = qh_batch_offset(b, i);
    total += nsub[i];
  }
  int *d_nsub = nullptr, *d_soff = nullptr;
  int64_t *d_off = nullptr, *d_tot = nullptr, *d_lev = nullptr, *d_str = nullptr;
  double *d_out = nullptr;
  int rc = qh_alloc_t(&d_nsub, n);
  if (!rc) rc = qh_alloc_t(&d_soff, n);
  if (!rc) rc = qh_alloc_t(&d_off, n);
  if (!rc) rc = qh_alloc_t(&d_tot, n);
  if (!rc) rc = qh_alloc_t(&d_lev, total);
  if (!rc) rc = qh_alloc_t(&d_str, total);
  if (!rc) rc = qh_alloc_t(&d_out, total);
  if (!rc) {
    hipStream_t s = qh_stream();
    (void)hipMemcpyAsync(d_nsub, nsub, 4 * n, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_soff, soff.data(), 4 * n, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_off, off.data(), 8 * n, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_tot, tot.data(), 8 * n, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_lev, levels, 8 * total, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_str, strides, 8 * total, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_batch_pops, dim3(n), dim3(256), 0, s, (const void *)x->d, x->lo, d_off, d_tot, d_nsub,
                       d_soff, d_lev, d_str, d_out);
    hipError_t e = hipMemcpyAsync(pops, d_out, 8 * total, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      qh_set_error("qh_batch_populations: %s", hipGetErrorString(e));
      rc = QH_EHIP;
    }
  }
  qh_free_ptr(d_nsub);
  qh_free_ptr(d_soff);
  qh_free_ptr(d_off);
  qh_free_ptr(d_tot);
  qh_free_ptr(d_lev);
  qh_free_ptr(d_str);
  qh_free_ptr(d_out);
  return rc;
}

// ensemble average of the normalised per-trajectory expectation values
extern "C" int qh_ens_expect(const qh_vec *ens, int64_t M, int64_t tot, int nops, const qh_op *ops, int nlev,
                             const qh_level *levels, double *re, double *im) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(ens && re && im && ens->prec == QH_FP64 && M >= 1 && M <= (1 << 30) && ens->n == M * tot, QH_EINVAL,
             "qh_ens_expect: FP64 ensemble of M trajectories of dimension tot required");
  qh_op *dops = nullptr;
  qh_level *dlev = nullptr;
  double2 *out = nullptr;
  rc = upload_chain(nops, ops, nlev, levels, &dops, &dlev, sizeof(double2) * (size_t)M, (void **)&out);
  double2 *h = nullptr;
  if (!rc) rc = qh_scratch(0, nullptr, sizeof(double2) * (size_t)M, (void **)&h);
  if (!rc) {
    hipLaunchKernelGGL(k_expect_ens, dim3((unsigned)M), dim3(256), 0, qh_stream(), (const double2 *)ens->d, tot,
                       dops, nops, dlev, out);
    if (hipMemcpyAsync(h, out, sizeof(double2) * M, hipMemcpyDeviceToHost, qh_stream()) != hipSuccess ||
        hipStreamSynchronize(qh_stream()) != hipSuccess) {
      qh_set_error("qh_ens_expect: device copy failed");
      rc = QH_EHIP;
    } else {
      double sr = 0.0, si = 0.0;
      for (int64_t b = 0; b < M; b++) {
        sr += h[b].x;
        si += h[b].y;
      }
      *re = sr / (double)M;
      *im = si / (double)M;
    }
  }
  return rc;
}

// ------------------------------------------- expectation without allocation
// qh_expect_dm_noalloc: the same reduction as qh_expect_dm for callers that may
// not allocate device memory.  Its op chain, level table, partial sums and total
// live in the file-scope array `scratch` below (each device holds its own copy
// until the process ends).  Host code never passes `scratch` itself to a kernel
// or copy: its address on the current device comes from hipGetSymbolAddress,
// resolved once per device and kept in g_scr.  Nothing in the array has to start
// at zero: each call writes the tables, the nb partial sums and the total it
// reads.  Every call runs on the library stream and waits for it before it
// returns, so calls from one host thread follow each other; calls that overlap
// on one GPU (several host threads at once) race on the array.
namespace {
const int NA_MAX_OPS = 64;      // operators in the chain
const int NA_MAX_LEV = 4096;    // entries of the level table
const size_t NA_OPS_B = 1024;   // 64 x 16 B (qh_op), 256-aligned
const size_t NA_LEV_B = 98304;  // 4096 x 24 B (qh_level), 256-aligned
const size_t NA_PART_B = sizeof(double2) * NB;
const size_t NA_BYTES = NA_OPS_B + NA_LEV_B + NA_PART_B + 256;
static_assert(sizeof(qh_op) * NA_MAX_OPS <= NA_OPS_B, "op area");
static_assert(sizeof(qh_level) * NA_MAX_LEV <= NA_LEV_B, "level area");
const int NA_MAX_DEV = 64;

__device__ __align__(256) unsigned char scratch[NA_BYTES];

void *g_scr[NA_MAX_DEV];
std::mutex g_scr_mu;

// the partial sums in index order, as the host loop of finish2 adds them, so
// that the total is bit-identical to qh_expect_dm's
__global__ void k_sum_partials_seq(const double2 *part, int nb, double2 *out) {
  double a = 0.0, b = 0.0;
  for (int i = 0; i < nb; i++) {
    a += part[i].x;
    b += part[i].y;
  }
  out[0] = make_double2(a, b);
}
}  // namespace

// device address of `scratch` on the current device (resolved on the first call
// per device; qh_init makes that call for its device)
int qh_noalloc_setup(char **out) {
  int dev = 0;
  QH_HIP(hipGetDevice(&dev));
  QH_REQUIRE(dev >= 0 && dev < NA_MAX_DEV, QH_ENODEV, "qh_expect_dm_noalloc: device %d beyond %d", dev,
             NA_MAX_DEV);
  std::lock_guard<std::mutex> lock(g_scr_mu);
  if (!g_scr[dev]) {
    void *p = nullptr;
    QH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QH_REQUIRE(p != nullptr, QH_EHIP, "hipGetSymbolAddress(scratch) returned no address on device %d", dev);
    g_scr[dev] = p;
  }
  if (out) *out = (char *)g_scr[dev];
  return QH_OK;
}

extern "C" int qh_expect_dm_noalloc(const qh_vec *x, int64_t tot, int nops, const qh_op *ops, int nlev,
                                    const qh_level *levels, double *re, double *im) {
  QH_REQUIRE(x && re && im && (nops == 0 || ops) && (nlev == 0 || levels), QH_EINVAL,
             "qh_expect_dm_noalloc: null argument");
  QH_REQUIRE(tot > 0 && tot * tot == x->n, QH_EINVAL, "qh_expect_dm_noalloc: vector size mismatch");
  QH_REQUIRE(nops >= 0 && nops <= NA_MAX_OPS, QH_ERANGE,
             "qh_expect_dm_noalloc: %d operators in the chain, at most %d", nops, NA_MAX_OPS);
  QH_REQUIRE(nlev >= 0 && nlev <= NA_MAX_LEV, QH_ERANGE,
             "qh_expect_dm_noalloc: %d level-table entries, at most %d", nlev, NA_MAX_LEV);
  for (int k = 0; k < nops; k++)
    QH_REQUIRE(ops[k].levels > 0 && ops[k].stride > 0 && ops[k].tab >= 0 && ops[k].tab + ops[k].levels <= nlev,
               QH_EINVAL, "observable op %d out of range", k);
  int rc = qh_ensure_init();
  if (rc) return rc;
  char *d = nullptr;
  rc = qh_noalloc_setup(&d);
  if (rc) return rc;
  qh_op *dops = (qh_op *)d;
  qh_level *dlev = (qh_level *)(d + NA_OPS_B);
  double2 *part = (double2 *)(d + NA_OPS_B + NA_LEV_B);
  double2 *total = (double2 *)(d + NA_OPS_B + NA_LEV_B + NA_PART_B);
  hipStream_t s = qh_stream();
  if (nops) QH_HIP(hipMemcpyAsync(dops, ops, sizeof(qh_op) * nops, hipMemcpyHostToDevice, s));
  if (nlev) QH_HIP(hipMemcpyAsync(dlev, levels, sizeof(qh_level) * nlev, hipMemcpyHostToDevice, s));
  const int nb = blocks_for(tot);
  hipLaunchKernelGGL(k_expect_dm, dim3(nb), dim3(256), 0, s, (const void *)x->d, x->lo, tot, dops, nops, dlev, part);
  hipLaunchKernelGGL(k_sum_partials_seq, dim3(1), dim3(1), 0, s, (const double2 *)part, nb, total);
  QH_HIP(hipGetLastError());
  double2 h;
  QH_HIP(hipMemcpyAsync(&h, total, sizeof(double2), hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  *re = h.x;
  *im = h.y;
  return QH_OK;
}
